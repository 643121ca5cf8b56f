// kernels.h -- device-side data layout shared by the kernels and the host
// driver in pcagpu.cpp.  Compiled only by hipcc.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/pcagpu.h"
#include "isa.h"

namespace pca {

constexpr int MAXT = PCAGPU_MAX_TESTS;
constexpr int MAXLEV = PCAGPU_MAX_KDEPTH + 1;   // levels 0..8
constexpr int LMAX = 3776;                      // == CAND_LMAX
constexpr int SCAN_WAVES = 4;                   // waves per workgroup
constexpr int SCAN_WG = SCAN_WAVES * 64;

// Per-task constants, passed by value (lands in the kernel-argument
// segment, read through the scalar cache).
struct DevTask {
    int T;
    int n_in;
    int max_cycles;          // clamped to 255
    int pad;
    uint16_t tgt[MAXT];      // expected output port value per test
    // per test, tgt | ldi_pass << 16 (bit i of ldi_pass: the i-th LDI
    // immediate of the grammar equals tgt), one dword for a scalar load
    uint32_t tgtw[MAXT];
    TestState root[MAXT];    // per-test state after the IN scaffold
};

// One device hit: a (prefix, suffix) whose state has a register equal to the
// target on every test.  flags bit 0: the program cannot pass (a test
// faulted or ran out of cycles) -- it still takes part in exact OEP.
struct HitRec {
    uint32_t prefix;         // index in the batch
    uint8_t mask;            // registers r1..r7 matching the target (any liveness)
    uint8_t flags;
    uint16_t j[PCAGPU_MAX_KDEPTH];   // candidate-list index at positions s..D-1
    uint16_t pad;
};
static_assert(sizeof(HitRec) == 24, "HitRec layout");

constexpr int VGMAX = 640;                      // == VG_MAX
constexpr int NBKINDS = 13;                     // == NBK

// The branch-mode grammar tables (CandTable) in one device block, shared by
// every handle on a device: per live mask m over r1..r7 the non-BR list in
// reference order, its count and BR split point, the total of live
// out-registers, and the last-position value groups as wave bit masks.
struct GrammarDev {
    uint16_t lists[128][LMAX];
    int count[128];
    int split[128];
    uint32_t cand_total[128];
    uint8_t vg_rdmask[128][VGMAX];            // rd values of a value group
    uint16_t vg_j[128][VGMAX][8];             // list index of (group, rd)
    uint64_t bits_init[128][NBKINDS];         // groups present per kind (ADD..ASR, ADDI, LDI)
    uint16_t bits_gidx[128][NBKINDS][64];     // group of a bit (0xFFFF: none)
    uint16_t kind_bits[128][64];              // per lane (bit position), the kinds with a group there
};

// What the kernels see: the shared tables, and the per-mask scalars, which a
// kernel may point at its own LDS copies.
struct CandDev {
    const GrammarDev *g;
    const int *count;           // [128]
    const int *split;           // [128]
    const uint32_t *cand_total; // [128]
    const uint64_t *bits_init;  // [128][NBKINDS]
};

// Work units of a scan launch.  k <= 1: unit = prefix.  k >= 2: units are
// grouped; a group is a prefix (k == 2, unit = position-s index) or a
// (prefix, position-s index) pair (k >= 3, unit = position s+1 index), so
// every unit is one parent state of roughly equal cost.
struct ScanGroup {
    uint32_t prefix;
    uint16_t js;
    uint16_t pad;
};

// Work unit of a scan launch for k >= 2 suffix positions: up to block_kids
// (<= 64) consecutive children of one node a level above the scanned
// parents' first level -- children at position s of prefix `prefix`
// (k == 2), or at position s+1 of the prefix followed by its position-s
// candidate js (k >= 3).  The block holds children j0 .. min(j0 +
// block_kids, count) - 1.
struct ScanBlock {
    uint32_t prefix;
    uint16_t js;
    uint16_t j0;
};
constexpr int BLOCK_KIDS_MAX = 64;
// a block of a shared-tile launch of 2- or 3-instruction suffixes may hold
// up to 128 children (its node scanned once for all of them)
constexpr int BLOCK_KIDS_SHARED = 128;

// Tile choice of a scan launch (picked by the host autotuner): children per
// block, tests run on every value group before survivors go to a queue,
// and how a 2- or 3-instruction suffix scans the last position.
struct ScanTile {
    uint16_t block_kids;   // 1..64 (1..128 for shared 2- and 3-instruction suffixes)
    uint8_t mask_tests;    // 1..6
    // 0: shared per block (value groups that do not read the register a
    // child writes are decided once on the block's node); 1: every child
    // scanned on its own
    uint8_t mode;
};
constexpr uint8_t TILE_SHARED = 0, TILE_PER_CHILD = 1;

struct ScanArgs {
    const pca_prefix_t *prefixes;
    uint32_t n_prefix;
    const ScanBlock *blocks;     // k >= 2: one unit per block; k <= 1: one unit per prefix
    uint64_t n_units;
    int D, s, k;
    ScanTile tile;
    CandDev cand;
    HitRec *hits;
    uint32_t hit_cap;
    uint32_t *hit_count;
    unsigned long long *counters;   // [0] candidates [1] suffixes [2] tests [3] parents
    unsigned long long *work;       // persistent work counter
};

// A level-p node whose state (registers and flags on every test) hashes to a
// query state; the host confirms it byte for byte.
struct MatchRec {
    uint32_t prefix;
    uint32_t query;
    uint16_t j[PCAGPU_MAX_KDEPTH];
};
static_assert(sizeof(MatchRec) == 24, "MatchRec layout");

struct MatchArgs {
    const pca_prefix_t *prefixes;
    uint32_t n_prefix;
    const ScanGroup *groups;
    const uint64_t *group_off;
    uint32_t n_groups;
    uint64_t n_units;
    int D, s, kk;                        // nodes at position p = s + kk - 1
    CandDev cand;
    const unsigned long long *q0;        // open-addressed set of test-0 row hashes
    uint64_t q0_mask;
    const unsigned long long *qf;        // [2 * cap] full-state hash, query id
    uint64_t qf_mask;
    MatchRec *out;
    uint32_t cap;
    uint32_t *count;
    unsigned long long *work;
};

void launch_match(const DevTask &task, const MatchArgs &a, int grid, hipStream_t st);

// Frontier (wavefront) expansion arguments.
struct ExpandArgs {
    const pca_prefix_t *parents;
    uint64_t n_parents;
    const uint64_t *child_off;       // [n_parents + 1] exclusive scan of child counts
    int D, pos;                      // children are placed at kernel position pos
    CandDev cand;
    unsigned long long *table;       // [2 * cap]: key, min child index
    uint64_t *child_hash;            // [children] each non-BR child's key, written by insert
    TestState *parent_rows;          // [n_parents * MAXT] each parent's test rows, written by insert
    uint64_t table_mask;             // cap - 1
    uint32_t *keep_bits;             // one bit per child
    uint64_t *keep_count;            // [n_parents] survivors per parent
    const uint64_t *keep_off;        // [n_parents + 1] exclusive scan of keep_count
    pca_prefix_t *out;
    unsigned long long *counters;    // [0] collisions [1] probe overflow
    unsigned long long *work;
    // insert / survive: each parent's child list is cut into `slices` runs
    // of whole 64-child blocks, one work unit each (a frontier of few
    // parents still fills the chip); a wave takes `grab` consecutive units
    // per atomic on `work` (0 or 1: one)
    uint32_t slices;
    uint32_t grab;
    // the children's keys are ANDed with this (all ones; fewer bits only to
    // test that colliding keys are caught)
    uint64_t key_mask;
};

void launch_alu_digest(uint32_t op, uint32_t b_lo, uint32_t b_hi, unsigned long long *out, hipStream_t st);
// One work item of a scan launch: a task (its resident record), a depth and
// split, a batch of its prefixes with their group table, and its own hit
// buffer.  Items of different tasks share one persistent launch.
struct ScanItem {
    const DevTask *task;
    const pca_prefix_t *prefixes;
    const ScanBlock *blocks;
    uint32_t n_prefix;
    uint32_t pad0;
    uint64_t unit_begin;             // first global unit of this item
    uint64_t n_units;
    int D, s, k;
    ScanTile tile;
    HitRec *hits;
    uint32_t *hit_count;
    uint32_t hit_cap;
    uint32_t pad2;
};

struct ScanLaunch {
    const ScanItem *items;
    uint32_t n_items;
    uint32_t levels;                 // state levels per wave in LDS: max over items of max(k, 1)
    uint64_t n_units;                // sum over the items
    CandDev cand;
    unsigned long long *counters;    // [0] candidates [1] suffixes [2] tests [3] parents
    unsigned long long *work;
    // one-item launches: per device test position, [2t] candidates that
    // reached it in the survivor phase and [2t+1] those it ended (null: off)
    unsigned long long *tstat;
    // consecutive work units a wave takes per atomic on `work` (0 or 1: one);
    // more for launches of many cheap units, whose single work counter
    // would otherwise serialise the waves
    uint32_t grab;
    uint32_t pad;
    // null, or [0] group tests credited to a block's children from their
    // node's shared scan, [1] group tests that shared scan ran
    unsigned long long *xcounters;
};

// Kernel-node form of a one-item scan launch, for a graph-captured flush:
// the argument block the node's kernelParams point into (the graph calls
// copy it), and the node parameters.
struct ScanNodeArgs {
    ScanLaunch l;
    ScanArgs a1;
    DevTask task;
    void *ptrs[3];
};
void scan_node_params(const ScanLaunch &l, const ScanItem &it, const DevTask &task, int grid, ScanNodeArgs *args,
                      hipKernelNodeParams *p);
// node that zeroes four word ranges p[i][0..n[i]) (32-bit words, n <= 256
// each); the graph's replacement for memset nodes
struct ZeroNodeArgs {
    uint32_t *p0, *p1, *p2, *p3;
    uint32_t n0, n1, n2, n3;
    void *ptrs[8];
};
void zero_node_params(uint32_t *const ptr[4], const uint32_t n[4], ZeroNodeArgs *args, hipKernelNodeParams *p);

void launch_table_init(unsigned long long *table, uint64_t cap, hipStream_t st);
// The frontier and match kernels, for their own occupancy queries.
enum class AuxKernel : int { ExpandInsert = 0, ExpandSurvive, ExpandCompact, Match, Count };
hipError_t aux_occupancy(AuxKernel which, int *blocks_per_cu);
// LDS bytes of a scan launch with `levels` state levels per wave
size_t scan_dyn_lds(uint32_t levels);
hipError_t scan_occupancy(bool multi, uint32_t levels, int *blocks_per_cu);
// host_items: the same items as l.items, in host memory; task0: the task
// record of item 0.  A one-item launch passes both in the kernel arguments.
void launch_scan(const ScanLaunch &l, const ScanItem *host_items, const DevTask &task0, int grid,
                 hipStream_t st);
void launch_expand_count(const ExpandArgs &a, uint64_t *counts, hipStream_t st);
void launch_expand_insert(const DevTask &task, const ExpandArgs &a, int grid, hipStream_t st);
void launch_expand_survive(const DevTask &task, const ExpandArgs &a, int grid, hipStream_t st);
void launch_expand_compact(const ExpandArgs &a, int grid, hipStream_t st);
// exclusive scan in place over n+1 u64 values (the last one receives the total)
void launch_exclusive_scan(uint64_t *data, uint64_t n, uint64_t *tmp, hipStream_t st);
uint64_t exclusive_scan_tmp_words(uint64_t n);

}  // namespace pca
