// host_core.h -- host side of the search path: task loading, the full VM
// used to re-verify every device hit, the disassembler, and the ordered
// candidate grammar.
#pragma once
#include <stdint.h>
#include <stddef.h>
#include <vector>

#include "../../include/pcagpu.h"
#include "isa.h"

namespace pca {

// Full VM (src/vm.c:14-217) for hit re-verification: memory, call stack and
// I/O ports included, so any program -- not only grammar ones -- runs as on
// the reference CPU VM.
struct Vm {
    uint16_t r[8];
    uint16_t mem[256];
    uint16_t prog[256];
    uint8_t pc, sp;
    uint16_t stack[16];
    uint8_t flags, status;
    uint64_t cycles;
    uint16_t ports[8];
};
enum { VM_RUNNING = 0, VM_HALTED = 1, VM_TIMEOUT = 2, VM_FAULT = 3 };

void vm_init(Vm &vm);
void vm_step(Vm &vm);
int vm_run(Vm &vm, uint64_t max_cycles);

int task_parse(const char *json, pca_task_t *t);
int task_load(const char *path, pca_task_t *t);
int task_evaluate(const pca_task_t *t, const uint16_t *prog, int len);
void disasm(uint16_t insn, char *buf, size_t len);

// Ordered branch-mode grammar at one kernel position (tools/enumerate.c:206-267).
int gen_candidates(uint16_t *out, int max, uint32_t live, int pos, int kernel_depth);
// search_gen_candidates_branchless (src/search_state.c:280-330)
int gen_candidates_branchless(uint16_t *out, int max, uint32_t live);

// The grammar factored for arithmetic indexing on the device: for every
// live mask m over r1..r7 (live = m << 1) the non-BR list in reference
// order, the split point `a` (ALU + CMP count) where the BR block goes, and
// the count.  Position p of a depth-D tree holds list(m)[0:a] ++
// BR(cond 1..6 x off 1..mf) ++ list(m)[a:], mf = D - p - 1.
constexpr int CAND_LMAX = 3776;   // >= 3,745 (all registers live)
// Last-position "value groups": candidates that differ only in rd write the
// same value vector (op applied to the same sources), so the device tests
// each distinct (op, sources, immediate) once and keeps a mask of the rd
// values still possible.  CMP is left out (it writes no register).
constexpr int VG_MAX = 640;        // >= distinct value groups of any mask
// Bit layout of the value groups for the wave-mask scan: kind 0..10 = the ALU
// ops ADD..ASR with bit 8*rs + rt, kind 11 = ADDI with bit 8*rd + index of
// the immediate in {-1, 1, -2, 2, -3, 3}, kind 12 = LDI with bit = index of
// the immediate in {0, 1, 2, 3, 4, 5, 8, 16, 32, 64, 128, 255}.
constexpr int NBK = 13;
// LDI immediates of the grammar, in list order (tools/enumerate.c:198)
inline constexpr uint8_t kLdiImms[12] = {0, 1, 2, 3, 4, 5, 8, 16, 32, 64, 128, 255};
struct CandTable {
    std::vector<uint16_t> list;   // [128][CAND_LMAX]
    int n[128];
    int a[128];
    std::vector<uint16_t> vg_insn;   // [128][VG_MAX] representative insn (rd field = 0 unless ADDI)
    std::vector<uint8_t> vg_rdmask;  // [128][VG_MAX] rd values present in the list
    std::vector<uint16_t> vg_j;      // [128][VG_MAX][8] list index of (group, rd)
    int vg_n[128];
    uint32_t cand_total[128];        // sum over the list of live out-registers
    std::vector<uint64_t> bits_init;    // [128][NBK] groups present, one bit each
    std::vector<uint16_t> bits_gidx;    // [128][NBK][64] group index of a bit (0xFFFF: none)
    void build(bool branchless);
    int count(uint32_t live, int mf) const { return n[(live >> 1) & 127] + (mf > 0 ? 6 * mf : 0); }
    uint16_t at(uint32_t live, int mf, int j) const
    {
        int m = (live >> 1) & 127;
        if (j < a[m]) return list[(size_t)m * CAND_LMAX + j];
        int nbr = mf > 0 ? 6 * mf : 0;
        if (j < a[m] + nbr) {
            int q = j - a[m];
            return enc_i(OP_BR, 1 + q / mf, 1 + q % mf);
        }
        return list[(size_t)m * CAND_LMAX + j - nbr];
    }
};

// Test order on the device.  Tests whose inputs look random (about half the
// bits set, far from 0 and 0xFFFF) come first: few candidates map such an
// input to the target by accident, while inputs like 0 or 1 are matched by
// many constant-ish candidates.  Equal-target tests are interleaved
// round-robin so a near-constant candidate fails within two tests.  The
// order only decides when a lane stops testing, never which programs pass.
std::vector<int> device_test_order(const pca_task_t *t);

// The grammar tables, built once per process on first use (thread-safe) and
// never modified afterwards.
const CandTable &shared_grammar();
// The search-state grammar's tables (branchless: no DIV, MOD, CMP or BR),
// built and kept the same way; count/at are called with mf = 0.
const CandTable &shared_grammar_branchless();

// Per-test root state: the registers after the IN scaffold
// (tools/enumerate.c:598-604) computed by running it on the VM.
void root_states(const pca_task_t *t, TestState *out /* [num_tests] */);

// Replay state of (scaffold + kernel[0..len) + HLT), key bytes in the order
// the reference hashes them: per test r0..r7 (lo, hi) then flags.
void oep_key_from_states(const TestState *st, int T, uint8_t *key /* T*17 */);

}  // namespace pca
