// kernels.hip -- gfx950 kernels of the search path.
//
//   scan_kernel      the hot path: every suffix of length k below a batch of
//                    split-depth prefixes (gpu/hybrid.cu:163-275 suffix_kernel)
//   expand_*         exact-OEP wavefront expansion of one tree level
//                    (gpu/wavefront.cu:330-414 expand_kernel), min-rank winners
//   scan_u64_*       exclusive prefix sums built from wave-level scans
//
// Work is organised per wave64: a wave owns one parent state at a time (the
// per-test registers live in LDS, one row per test), lanes = candidates of the
// last kernel position, and each lane exits the test loop on its first
// mismatch; the wave leaves the loop when a ballot finds no lane still alive.
#include <cstring>

#include "kernels.h"

namespace pca {

namespace {

// two 16-bit lanes of one 32-bit register (packed 16-bit integer ops)
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));
typedef short i16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
// the wave's lanes where p holds, from the compare itself (no widening of
// the predicate to an integer and second compare, as __ballot(int) does)
__device__ __forceinline__ uint64_t bal(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ __forceinline__ bool any_lane(bool p) { return __builtin_amdgcn_ballot_w64(p) != 0ull; }
// lane 0's value, known uniform to the compiler (SGPRs, uniform branches)
__device__ __forceinline__ unsigned long long uni64(unsigned long long v)
{
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

__device__ __forceinline__ int cand_count(const CandDev &c, uint32_t live, int mf)
{
    return c.count[(live >> 1) & 127] + (mf > 0 ? 6 * mf : 0);
}

// The reference order at one position: ALU ++ CMP ++ BR(mf) ++ LDI ++ ADDI
// (tools/enumerate.c:206-267), indexed arithmetically.
__device__ __forceinline__ uint32_t cand_at(const CandDev &c, uint32_t live, int mf, int j)
{
    const int m = (live >> 1) & 127;
    const int a = c.split[m];
    if (j < a) return c.g->lists[m][j];
    const int nbr = mf > 0 ? 6 * mf : 0;
    if (j < a + nbr) {
        const int q = j - a;
        return enc_i(OP_BR, 1 + q / mf, 1 + q % mf);
    }
    return c.g->lists[m][j - nbr];
}

__device__ __forceinline__ unsigned long long wave_sum(unsigned long long v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Inclusive prefix sum over the wave's 64 lanes in six DPP adds: shifts by
// 1, 2, 4 and 8 lanes within each 16-lane row (lanes shifted in from outside
// the row read 0), then lane 15's row total added to rows 1 and 3 and lane
// 31's running total to rows 2 and 3.  Every lane must be active.
__device__ __forceinline__ uint32_t wave_incl_scan32(uint32_t v)
{
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);   // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xF, 0xF, true);   // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xF, 0xF, true);   // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xF, 0xF, true);   // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xA, 0xF, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xC, 0xF, false);  // row_bcast:31
    return v;
}

// Replays a prefix from the root into an LDS level (lanes = tests).  The
// state is stepped in place in LDS so that the register index chosen by the
// instruction addresses LDS, not a private array.
__device__ __forceinline__ PackedState load_ps(const TestState *t)
{
    const uint32_t *q = reinterpret_cast<const uint32_t *>(t);
    PackedState s;
#pragma unroll
    for (int i = 0; i < 5; i++) s.w[i] = q[i];
    return s;
}

__device__ __forceinline__ void store_ps(TestState *t, const PackedState &s)
{
    uint32_t *q = reinterpret_cast<uint32_t *>(t);
#pragma unroll
    for (int i = 0; i < 5; i++) q[i] = s.w[i];
}

// Replays a prefix from the root into an LDS level (lanes = tests), the
// state held in registers (PackedState) while it steps.
__device__ __forceinline__ void replay_prefix(const DevTask &task, const pca_prefix_t *p, int n,
                                              TestState *dst, int lane)
{
    if (lane < task.T) {
        PackedState st = load_ps(&task.root[lane]);
        for (int q = 0; q < n; q++) step_packed(st, p->insns[q], (uint32_t)q);
        store_ps(&dst[lane], st);
    }
}

// One level down (lanes = tests): dst = src stepped by the wave-uniform insn.
__device__ __forceinline__ void step_level(const DevTask &task, const TestState *src, TestState *dst,
                                           uint32_t insn, int pos, int lane)
{
    if (lane < task.T) {
        PackedState st = load_ps(&src[lane]);
        step_packed(st, insn, (uint32_t)pos);
        store_ps(&dst[lane], st);
    }
}

// The wave's work counters, a row in LDS that lane 0 adds to (no registers
// held for them): [0] candidates [1] suffixes [2] tests [3] parents.
struct WaveCounters {
    unsigned long long *c;
};

// Value a last-position candidate of opcode OP writes (flags are not needed
// there: only OUT r; HLT follow).
template <uint32_t OP>
__device__ __forceinline__ uint32_t op_value(uint32_t x, uint32_t y, uint32_t imm)
{
    if (OP == OP_ADD) return (x + y) & 0xFFFFu;
    if (OP == OP_SUB) return (x - y) & 0xFFFFu;
    if (OP == OP_MUL) return (x * y) & 0xFFFFu;
    if (OP == OP_DIV) return sdiv16(x, y);
    if (OP == OP_MOD) return smod16(x, y);
    if (OP == OP_AND) return x & y;
    if (OP == OP_OR) return x | y;
    if (OP == OP_XOR) return x ^ y;
    if (OP == OP_SHL) return (x << (y & 15u)) & 0xFFFFu;
    if (OP == OP_SHR) return x >> (y & 15u);
    if (OP == OP_ASR) return (uint32_t)((int32_t)(int16_t)(uint16_t)x >> (y & 15u)) & 0xFFFFu;
    if (OP == OP_LDI) return imm;
    if (OP == OP_ADDI) return (x + ((uint32_t)(int32_t)(int8_t)(uint8_t)imm)) & 0xFFFFu;
    return 0;
}

// Test loop of the lanes whose candidate has opcode OP (`mine`).  A lane
// stays in while its rd still equals the target on every test so far (or,
// for DIV/MOD with another matching register, until a fault is ruled out);
// the wave leaves as soon as a ballot finds no lane in.
template <uint32_t OP>
__device__ __forceinline__ void test_loop(const DevTask &task, const TestState *P, uint64_t skipm, bool mine,
                                          uint32_t rd, uint32_t sa, uint32_t rt, uint32_t imm, uint32_t other,
                                          bool &rd_ok, bool &fault, uint32_t &tests)
{
    constexpr bool DIVMOD = OP == OP_DIV || OP == OP_MOD;
    constexpr bool TWO = OP != OP_LDI && OP != OP_ADDI;
    const int T = task.T;
    bool need = mine;
    for (int t = 0; t < T; t++) {
        if (!any_lane(need)) break;
        if (need) {
            uint32_t v;
            if ((skipm >> t) & 1ull) {
                v = P[t].r[rd];
            } else {
                const uint32_t x = OP == OP_LDI ? 0u : P[t].r[sa];
                const uint32_t y = TWO ? P[t].r[rt] : 0u;
                if (DIVMOD && y == 0u) {
                    fault = true;      // the VM stops: rd keeps its value
                    v = P[t].r[rd];
                } else {
                    v = op_value<OP>(x, y, imm);
                }
            }
            rd_ok = rd_ok && v == task.tgt[t];
            tests++;
            // rd must be checked on every test (a faulted test keeps the old
            // rd); with another matching register, a DIV/MOD lane also stays
            // until a fault is ruled out, since a fault makes it unpassable
            need = rd_ok || (DIVMOD && other != 0u && !fault);
        }
    }
}

// Stage A of a chunk: tests 0 and 1 for the lanes of opcode OP, straight
// line (no early exit), four LDS reads in flight together.
template <uint32_t OP>
__device__ __forceinline__ void pair_eval(const DevTask &task, const TestState *P, uint64_t skipm, bool mine,
                                          uint32_t rd, uint32_t sa, uint32_t rt, uint32_t imm, bool &rd_ok)
{
    constexpr bool DIVMOD = OP == OP_DIV || OP == OP_MOD;
    constexpr bool TWO = OP != OP_LDI && OP != OP_ADDI;
    if (!mine) return;
    uint32_t v[2];
#pragma unroll
    for (int t = 0; t < 2; t++) {
        const uint32_t old = P[t].r[rd];
        const uint32_t x = OP == OP_LDI ? 0u : P[t].r[sa];
        const uint32_t y = TWO ? P[t].r[rt] : 0u;
        uint32_t r = (DIVMOD && y == 0u) ? old : op_value<OP>(x, DIVMOD && y == 0u ? 1u : y, imm);
        v[t] = ((skipm >> t) & 1ull) ? old : r;
    }
    rd_ok = rd_ok && v[0] == task.tgt[0] && v[1] == task.tgt[1];
}

// What the last-position scan below one parent state P (kernel position
// D-1) needs to know about P; hits are last-position candidates whose state
// has a target-matching register.
struct ParentSummary {
    uint32_t pm;       // registers r1..r7 equal to the target on every test
    uint64_t skipm;    // tests on which the last instruction does not execute
    uint32_t dead;     // some test already faulted or cannot fit max_cycles
};

// Lanes = tests.  my_tgt: this lane's target (lane = test), held in a
// register by the caller; st: this lane's parent state (valid when act).
// Each register half is compared with the target by a 16-bit test of the
// XOR (low half) or its magnitude (high half); the ballots are the summary.
__device__ __forceinline__ ParentSummary summarize_state(const DevTask &task, const PackedState &st, bool act,
                                                         uint32_t pos, uint32_t my_tgt)
{
    // every ballot is of one compare; the active lanes are ANDed in as a
    // mask (a ballot of a combined predicate makes the compiler rebuild it
    // in a vector register and compare again)
    const uint64_t am = bal(act);
    const uint32_t tt = my_tgt | (my_tgt << 16);
    ParentSummary s;
    uint32_t pm = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t d = st.w[i] ^ tt;
        if (i > 0) pm |= ((bal((d & 0xFFFFu) != 0u) & am) == 0ull ? 1u : 0u) << (2 * i);
        pm |= ((bal(d > 0xFFFFu) & am) == 0ull ? 1u : 0u) << (2 * i + 1);
    }
    s.pm = pm;
    // skipped by a pending BR (resume past pos) or the VM already stopped
    // (frozen, the byte above resume: any nonzero value exceeds pos)
    const uint32_t w4 = st.w[4], ex = w4 >> 24, mc = (uint32_t)task.max_cycles;
    const uint64_t nm = bal(((w4 >> 8) & 0xFFFFu) > pos) & am;
    s.skipm = nm;
    // out of cycles: the last instruction, OUT and HLT (OUT and HLT alone
    // on a skipped test) would pass max_cycles
    const uint64_t cyc = bal(ex + 2u > mc) | (bal(ex + 3u > mc) & ~nm);
    s.dead = ((bal((w4 & 0xFF0000u) != 0u) | cyc) & am) != 0ull ? 1u : 0u;
    return s;
}

__device__ __forceinline__ ParentSummary summarize_parent(const DevTask &task, const TestState *P, uint32_t pos,
                                                          int lane, uint32_t my_tgt)
{
    const bool act = lane < task.T;
    PackedState st = {};
    if (act) st = load_ps(&P[lane]);
    return summarize_state(task, st, act, pos, my_tgt);
}

// Registers r1..r7 (bits 1..7) of one test's state row equal to that test's
// target tg; read from the row when a scan needs it (a test the last
// instruction skips, a zero divisor), not kept per parent.
__device__ __forceinline__ uint32_t row_regmatch(const TestState *row, uint32_t tg)
{
    const uint32_t *q = reinterpret_cast<const uint32_t *>(row);
    const uint32_t t2 = tg * 0x10001u;
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t d = q[i] ^ t2;
        m |= ((d & 0xFFFFu) == 0u ? 1u : 0u) << (2 * i);
        m |= (d <= 0xFFFFu ? 1u : 0u) << (2 * i + 1);
    }
    return m & 0xFEu;
}

// The same for a row every lane reads (a test position the whole wave is
// at): its words read once into scalar registers, the compares scalar.
__device__ __forceinline__ uint32_t row_regmatch_uni(const TestState *row, uint32_t tg)
{
    const uint32_t *q = reinterpret_cast<const uint32_t *>(row);
    const uint32_t t2 = tg * 0x10001u;
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t d = uni(q[i]) ^ t2;
        m |= ((d & 0xFFFFu) == 0u ? 1u : 0u) << (2 * i);
        m |= (d <= 0xFFFFu ? 1u : 0u) << (2 * i + 1);
    }
    return m & 0xFEu;
}

// The candidate indices of a scanned parent at positions s..D-2: a parent at
// most two positions below the split keeps them in registers (j0, j1),
// deeper ones are read from the wave's LDS row.
struct JPath {
    uint32_t j0, j1;
    const uint32_t *lds;   // null: j0, j1
};
__device__ __forceinline__ uint32_t jp_at(const JPath &p, int q) { return p.lds ? p.lds[q] : (q == 0 ? p.j0 : p.j1); }

// Pooled survivor phase: parents whose surviving value groups share one
// 64-lane drain.  Each wave keeps NSLOT parent states in LDS slots, with
// what a drain needs of each parent.
constexpr uint32_t NSLOT = 4;
// Everything a wave keeps in LDS besides the state levels above the scanned
// parents, in one block (one base address, fields at fixed offsets).
struct WaveLds {
    TestState slot[NSLOT][MAXT];          // the scanned parents' states
    uint32_t queue[128];                  // pooled survivors
    uint32_t cq[128];                     // scan_parent's candidates
    uint32_t live[MAXLEV], idx[MAXLEV];   // the suffix walk: live masks, candidate indices
    unsigned long long skip[NSLOT];       // per slot, tests the last instruction does not run on
    uint32_t meta[NSLOT][2];              // per slot, prefix index; live-mask index | dead << 8
    uint32_t path[NSLOT][MAXLEV];         // per slot, candidate indices at positions s..D-2
    unsigned long long wc[4];             // work counters (WaveCounters)
    unsigned long long wx[2];             // shared-scan counters (ScanLaunch::xcounters)
    uint32_t evl[64];                     // per lane, phase-A group evaluations
    uint32_t tst[MAXT][2];                // survivor-phase test statistics
    // per lane, phase A at test position 2 (the only reordered position it
    // runs, at most 3 tests): groups that reached it (low half), groups it
    // ended (high half); folded into tst[2] when the wave leaves
    unsigned long long pa2[64];
};
struct Pool {                      // wave-uniform bookkeeping
    uint32_t qn;                   // entries queued
    uint32_t pend;                 // entries queued per slot, 8 bits each
    uint32_t cur;                  // slot of the parent being scanned
    uint32_t nd;                   // drains so far
    uint32_t ns;                   // slots in use: NSLOT, or 2 when slots 2..3 hold ShareLds
};

// The record is written field by field (no private copy).
__device__ __forceinline__ void emit_hit(const ScanArgs &a, int k, uint32_t prefix_idx, const JPath &jpath,
                                         uint32_t jlast, uint32_t mask, uint32_t flags)
{
    const uint32_t slot = atomicAdd(a.hit_count, 1u);
    if (slot < a.hit_cap) {
        HitRec *h = a.hits + slot;
        h->prefix = prefix_idx;
        h->mask = (uint8_t)mask;
        h->flags = (uint8_t)flags;
        h->pad = 0;
        for (int q = 0; q < PCAGPU_MAX_KDEPTH; q++)
            h->j[q] = (uint16_t)(q + 1 < k ? jp_at(jpath, q) : (q + 1 == k ? jlast : 0u));
    }
}

// Survivor phase of the pooled value groups: queue entries [0, cnt), one
// per lane, each tested from test t0 (where phase A stopped) to the last
// with an early exit, against its own parent's state (its slot).  A group
// still matching on every test gives one hit per allowed rd, with its
// parent's prefix and suffix path.  A test that some entry's parent skips
// (pending BR, VM stopped) keeps that entry's rd values whose register
// already equals the target.
__device__ __forceinline__ void pool_drain(const DevTask &task, const ScanArgs &a, int k, const CandDev &cd,
                                           WaveLds &sv, uint32_t cnt, int t0, int lane, WaveCounters &wc,
                                           uint32_t (*tst)[2], Pool &pool)
{
    const int T = task.T;
    const bool valid = (uint32_t)lane < cnt;
    const uint32_t e = valid ? sv.queue[lane] : 0u;
    const uint32_t q = e & 15u, bit = (e >> 4) & 63u, s = (e >> 19) & (NSLOT - 1u);
    uint32_t ok = valid ? (e >> 10) & 0xFFu : 0u;
    bool flt = (e >> 18) & 1u;
    // tests skipped by any entry's parent (uniform), entries left per slot
    uint64_t skor = 0;
#pragma unroll
    for (uint32_t z = 0; z < NSLOT; z++) {
        const uint64_t b = bal(valid && s == z);
        if (b) skor |= uni64(sv.skip[z]);
        pool.pend -= (uint32_t)__popcll(b) << (8u * z);
    }
    const uint64_t skl = skor ? sv.skip[s] : 0ull;
    const TestState *P = sv.slot[s];
    // survivor-phase reach / end counts from one drain in eight are enough
    // to order the tests, at an eighth of their cost
    const bool tst8 = tst && (pool.nd & 7u) == 0u;
    pool.nd++;
    const uint32_t xa = bit >> 3, yb = bit & 7u, i = bit & 7u;
    const uint32_t aimm = (i & 1u) ? (i >> 1) + 1u : 0x10000u - ((i >> 1) + 1u);
    const bool divmod = q == OP_DIV - OP_ADD || q == OP_MOD - OP_ADD;
    bool any_div = any_lane(divmod && ok != 0u);
    uint32_t n_ev = 0;
    const uint32_t aimm2 = (aimm & 0xFFFFu) * 0x10001u;
    for (int t_ = t0; t_ < T; t_++) {
        int tt = (int)uni((uint32_t)t_);
        const uint64_t al = bal(ok != 0u);
        if (al == 0ull) break;
        const uint32_t before = (uint32_t)__popcll(al);
        n_ev += before;
        if (any_div) any_div = any_lane(divmod && ok != 0u);
        if (!any_div && tt + 1 < T && ((skor >> tt) & 3ull) == 0ull) {
            // tests tt and tt+1 at once: each 16-bit half of a register
            // holds one test's operand, and the packed 16-bit ops compute
            // both values (mod 2^16, as the VM does)
            const uint32_t tw0 = task.tgtw[tt], tw1 = task.tgtw[tt + 1];
            const uint32_t tg2 = (tw0 & 0xFFFFu) | (tw1 << 16);
            const uint32_t x2 = (uint32_t)P[tt].r[xa] | ((uint32_t)P[tt + 1].r[xa] << 16);
            const uint32_t y2 = (uint32_t)P[tt].r[yb] | ((uint32_t)P[tt + 1].r[yb] << 16);
            const u16x2 X = __builtin_bit_cast(u16x2, x2), Y = __builtin_bit_cast(u16x2, y2);
            const u16x2 S = Y & (unsigned short)15;
            uint32_t v = __builtin_bit_cast(uint32_t, X + Y);
            v = q == 1u ? __builtin_bit_cast(uint32_t, X - Y) : v;
            v = q == 2u ? __builtin_bit_cast(uint32_t, X * Y) : v;
            v = q == 5u ? (x2 & y2) : v;
            v = q == 6u ? (x2 | y2) : v;
            v = q == 7u ? (x2 ^ y2) : v;
            v = q == 8u ? __builtin_bit_cast(uint32_t, X << S) : v;
            v = q == 9u ? __builtin_bit_cast(uint32_t, X >> S) : v;
            v = q == 10u ? __builtin_bit_cast(uint32_t, __builtin_bit_cast(i16x2, x2) >> __builtin_bit_cast(i16x2, S))
                         : v;
            v = q == 11u ? __builtin_bit_cast(uint32_t, X + __builtin_bit_cast(u16x2, aimm2)) : v;
            const uint32_t d = v ^ tg2;
            bool m0 = (d & 0xFFFFu) == 0u, m1 = (d >> 16) == 0u;
            if (q == 12u) {
                m0 = (tw0 >> (16 + bit)) & 1u;
                m1 = (tw1 >> (16 + bit)) & 1u;
            }
            // the second test is evaluated for the lanes the first kept
            const uint32_t mid = (uint32_t)__popcll(bal(ok != 0u && m0));
            n_ev += mid;
            if (!(m0 && m1)) ok = 0;
            if (tst8) {
                const uint32_t after = (uint32_t)__popcll(bal(ok != 0u));
                if (lane == 0) {
                    tst[tt][0] += before;
                    tst[tt][1] += before - mid;
                    tst[tt + 1][0] += mid;
                    tst[tt + 1][1] += mid - after;
                }
            }
            t_ = tt + 1;
            continue;
        }
        const bool sk = (skl >> tt) & 1ull;
        const uint32_t tw = task.tgtw[tt];
        const uint32_t tg = tw & 0xFFFFu;
        const uint32_t x = P[tt].r[xa], y = P[tt].r[yb];
        const uint32_t sh = y & 15u;
        uint32_t v = x + y;
        v = q == 1u ? x - y : v;
        v = q == 2u ? x * y : v;
        v = q == 5u ? (x & y) : v;
        v = q == 6u ? (x | y) : v;
        v = q == 7u ? (x ^ y) : v;
        v = q == 8u ? x << sh : v;
        v = q == 9u ? x >> sh : v;
        v = q == 10u ? (uint32_t)((int32_t)(int16_t)(uint16_t)x >> sh) : v;
        v = q == 11u ? x + aimm : v;
        bool match = (v & 0xFFFFu) == tg;
        if (q == 12u) match = (tw >> (16 + bit)) & 1u;
        if (any_div && divmod && !sk) {
            if (y == 0u) {
                ok &= row_regmatch(&sv.slot[s][tt], tg);
                flt = true;
                match = true;
            } else {
                const int32_t sx = (int16_t)(uint16_t)x, sy = (int16_t)(uint16_t)y;
                const int32_t qt = sdiv16_q(sx, sy);
                match = ((uint32_t)(q == OP_DIV - OP_ADD ? qt : sx - qt * sy) & 0xFFFFu) == tg;
            }
        }
        if (sk)
            ok &= row_regmatch(&sv.slot[s][tt], tg);
        else if (!match)
            ok = 0;
        if (tst8) {
            const uint32_t after = (uint32_t)__popcll(bal(ok != 0u));
            if (lane == 0) {
                tst[tt][0] += before;
                tst[tt][1] += before - after;
            }
        }
    }
    if (lane == 0) wc.c[2] += n_ev;
    if (ok != 0u) {
        const uint32_t mt = sv.meta[s][1], mm = mt & 127u;
        const uint32_t g = cd.g->bits_gidx[mm][q][bit];
        const uint32_t okf = cd.g->vg_rdmask[mm][g] & ok;
        const uint32_t fl = ((mt >> 8) & 1u) | ((divmod && flt) ? 1u : 0u);
        JPath jp;
        jp.j0 = jp.j1 = 0;
        jp.lds = sv.path[s];
        const uint32_t pidx = sv.meta[s][0];
        for (uint32_t r = 1; r < 8; r++)
            if (okf >> r & 1u) emit_hit(a, k, pidx, jp, cd.g->vg_j[mm][g][r], 1u << r, fl);
    }
}

// Parent with no target-matching register (the common case): scan the last
// position by value groups.  Lane 8*a + b holds the source pair (a, b) and
// a mask (kl) with one bit per ALU kind whose (op, a, b) group exists for
// the live mask and still matches: one test of every group of the parent
// costs one evaluation and one compare per op across the wave, with no
// wave-uniform state per kind (no ballots until the survivors are queued).
// ADDI groups use lane 8*rd + i (source rd, i-th immediate of -1, 1, -2, 2,
// -3, 3); LDI groups are per-test constants (DevTask::tgtw, lane i the i-th
// immediate).  A test on which the last instruction does not run (skipped
// by a pending BR, VM stopped) keeps the rd values whose old register
// already equals the target: that regmatch mask is ANDed into the rd masks
// of all groups (rm_and) and prunes the groups left without an rd.  DIV and
// MOD keep a per-lane rd mask, since a zero divisor (a fault) makes a test
// behave like that for that lane only.  The per-lane kind masks of each
// live mask come from the grammar table (GrammarDev::kind_bits).
// The parent's state P sits in the wave's slot pool.cur; its surviving
// groups are queued for the pooled survivor phase (pool_drain).
__device__ __forceinline__ void scan_parent_bits(const DevTask &task, const ScanArgs &a, int k, const CandDev &cd,
                                                 const TestState *P,
                                                 uint32_t live, uint32_t prefix_idx, const JPath &jpath,
                                                 int lane, WaveCounters &wc,
                                                 WaveLds &sv, Pool &pool, uint32_t (*tst)[2],
                                                 const ParentSummary &ps)
{
    const int T = task.T;
    const int m = (live >> 1) & 127;
    if (lane == 0) {
        wc.c[0] += cd.cand_total[m];
        wc.c[1] += (unsigned)cd.count[m];
    }
    // per lane (source pair a, b; ADDI rd, i; LDI i), bit q: the group of
    // kind q on this lane is still matching (DIV and MOD: rd masks okd, okm)
    uint32_t kl = cd.g->kind_bits[m][lane];
    // every DIV / MOD group has all of r1..r7 as rd (the host checks this)
    uint32_t okd = (kl >> (OP_DIV - OP_ADD) & 1u) ? 0xFEu : 0u;
    uint32_t okm = (kl >> (OP_MOD - OP_ADD) & 1u) ? 0xFEu : 0u;
    kl &= ~((1u << (OP_DIV - OP_ADD)) | (1u << (OP_MOD - OP_ADD)));
    bool fault = false;
    uint32_t rm_and = 0xFFu;
    const uint32_t ra = (uint32_t)lane >> 3, rb = (uint32_t)lane & 7u;
    const uint32_t addi_imm = (rb & 1u) ? (rb >> 1) + 1u : 0x10000u - ((rb >> 1) + 1u);
    uint32_t ev = 0;   // this lane's group evaluations
    // phase A: the first tests, every group of the parent at once
    const int TA = T < (int)a.tile.mask_tests ? T : (int)a.tile.mask_tests;
    int t = 0;
    for (; t < TA; t++) {
        const uint32_t nal = (uint32_t)__popc(kl) + (okd != 0u ? 1u : 0u) + (okm != 0u ? 1u : 0u);
        if (!any_lane(nal != 0u)) break;
        // group evaluations of this test; once a skipped test has narrowed
        // the rd values (rm_and), groups left without one are not told
        // apart here, so such parents are not counted further
        if (rm_and == 0xFFu) ev += nal;
        if ((ps.skipm >> t) & 1ull) {
            const uint32_t rmt = row_regmatch_uni(&P[t], task.tgt[t]);
            okd &= rmt;
            okm &= rmt;
            rm_and &= rmt;
            continue;
        }
        const uint32_t x = P[t].r[ra], y = P[t].r[rb];
        const uint32_t tw = task.tgtw[t];
        const uint32_t tg = tw & 0xFFFFu;
        const uint32_t sh = y & 15u;
        uint32_t keep = (lane < 16 && ((tw >> (16 + lane)) & 1u)) ? 1u << 12 : 0u;
        keep |= (((x + y) & 0xFFFFu) == tg ? 1u : 0u) << 0;
        keep |= (((x - y) & 0xFFFFu) == tg ? 1u : 0u) << 1;
        keep |= (((x * y) & 0xFFFFu) == tg ? 1u : 0u) << 2;
        keep |= ((x & y) == tg ? 1u : 0u) << 5;
        keep |= ((x | y) == tg ? 1u : 0u) << 6;
        keep |= ((x ^ y) == tg ? 1u : 0u) << 7;
        keep |= (((x << sh) & 0xFFFFu) == tg ? 1u : 0u) << 8;
        keep |= ((x >> sh) == tg ? 1u : 0u) << 9;
        keep |= (((uint32_t)((int32_t)(int16_t)(uint16_t)x >> sh) & 0xFFFFu) == tg ? 1u : 0u) << 10;
        keep |= (((x + addi_imm) & 0xFFFFu) == tg ? 1u : 0u) << 11;
        kl &= keep;
        if (any_lane((okd | okm) != 0u)) {
            // a zero divisor stops the VM and rd keeps its value: the test
            // passes for the rd values whose register already matches (the
            // test's match mask, the same for every lane); no branch on it
            const bool z = y == 0u;
            const uint32_t rmt = any_lane(z) ? row_regmatch_uni(&P[t], tg) : 0u;
            const int32_t sx = (int16_t)(uint16_t)x, sy = z ? 1 : (int32_t)(int16_t)(uint16_t)y;
            const int32_t qt = sdiv16_q(sx, sy);
            const bool dq = ((uint32_t)qt & 0xFFFFu) == tg, dr = ((uint32_t)(sx - qt * sy) & 0xFFFFu) == tg;
            okd &= z ? rmt : (dq ? 0xFFu : 0u);
            okm &= z ? rmt : (dr ? 0xFFu : 0u);
            fault = fault || (z && (okd != 0u || okm != 0u));
        }
        if (tst && t == 2) {   // position 2 is reordered by these counts (phase A runs at most 3 tests)
            const uint32_t left = (uint32_t)__popc(kl) + (okd != 0u ? 1u : 0u) + (okm != 0u ? 1u : 0u);
            __hip_atomic_fetch_add(&sv.pa2[lane], (unsigned long long)nal | ((unsigned long long)(nal - left) << 32),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
    }
    if (ev) __hip_atomic_fetch_add(&sv.evl[lane], ev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    if (rm_and == 0u) return;   // no rd value is possible for any group
    // the kinds with a group still matching on this lane, and where this
    // lane's entries go among the parent's survivors (lane order, then kind)
    const uint32_t sm = kl | (okd != 0u ? 1u << (OP_DIV - OP_ADD) : 0u) | (okm != 0u ? 1u << (OP_MOD - OP_ADD) : 0u);
    const uint32_t c = (uint32_t)__popc(sm);
    const uint32_t incl = wave_incl_scan32(c);
    const uint32_t total = uni(__builtin_amdgcn_readlane(incl, 63));
    if (total == 0u) return;
    const uint32_t excl = incl - c;
    // phase B is pooled: the survivors wait in the wave's queue with those of
    // earlier parents (whose states stay in their slots) and are tested 64
    // at a time by pool_drain.  Queue entry: kind | bit << 4 | rd mask << 10
    // | fault << 18 | slot << 19.
    const uint32_t cur = pool.cur;
    if (lane == 0) {
        sv.meta[cur][0] = prefix_idx;
        sv.meta[cur][1] = (uint32_t)m | (ps.dead ? 0x100u : 0u);
        sv.skip[cur] = ps.skipm;
    }
    if (lane + 1 < k) sv.path[cur][lane] = jp_at(jpath, lane);
    const uint32_t next = cur + 1u == pool.ns ? 0u : cur + 1u;
    uint32_t pushed = 0;   // this parent's entries queued so far (the first `pushed` in lane order)
    for (;;) {
        if (pushed < total) {
            // as many entries as the queue has room for
            const uint32_t take = min(128u - pool.qn, total - pushed);
            if (c != 0u && excl < pushed + take && excl + c > pushed) {
                uint32_t mk = sm, g = excl;
                while (mk) {
                    const uint32_t q = (uint32_t)__builtin_ctz(mk);
                    mk &= mk - 1u;
                    if (g >= pushed && g < pushed + take) {
                        const bool dm = q == OP_DIV - OP_ADD || q == OP_MOD - OP_ADD;
                        const uint32_t okq = q == OP_DIV - OP_ADD ? okd : (dm ? okm : rm_and);
                        sv.queue[pool.qn + (g - pushed)] =
                            q | ((uint32_t)lane << 4) | (okq << 10) | ((dm && fault) ? 1u << 18 : 0u) | (cur << 19);
                    }
                    g++;
                }
            }
            pool.qn += take;
            pool.pend += take << (8u * cur);
            pushed += take;
        }
        const bool all = pushed == total;
        // drain 64 when that many wait; once this parent is queued, drain
        // the rest too if the next parent's slot still has entries
        uint32_t cnt = 0;
        if (pool.qn >= 64u)
            cnt = 64u;
        else if (all && ((pool.pend >> (8u * next)) & 0xFFu) != 0u)
            cnt = pool.qn;
        if (cnt) {
            wave_sync();
            pool_drain(task, a, k, cd, sv, cnt, TA, lane, wc, tst, pool);
            wave_sync();
            const uint32_t rest = pool.qn - cnt;
            if (cnt == 64u && (uint32_t)lane < rest) sv.queue[lane] = sv.queue[64 + lane];
            pool.qn = rest;
            wave_sync();
            continue;
        }
        if (all) break;
    }
    pool.cur = next;
}

// Parent where some register already equals the target: every candidate
// that leaves it alone is a hit, so go candidate by candidate.
__device__ __forceinline__ void scan_parent(const DevTask &task, const ScanArgs &a, int k, const CandDev &cd,
                                            const TestState *P,
                                            uint32_t live, uint32_t prefix_idx, const JPath &jpath, int lane,
                                            WaveCounters &wc, uint32_t *queue, const ParentSummary &ps)
{
    const int T = task.T;
    const uint32_t pm = ps.pm;
    const uint64_t skipm = ps.skipm;
    const uint32_t dead = ps.dead;

    const int m = (live >> 1) & 127;
    const int n = cd.count[m];
    const uint16_t *L = cd.g->lists[m];
    uint32_t n_cand = 0, n_suf = 0, n_tests = 0;
    uint32_t qn = 0;   // queued survivors (wave-uniform)
    const bool two = T >= 2;

    // full evaluation of queue entries [0, cnt): lanes = entries
    auto drain = [&](uint32_t cnt) {
        const bool valid = (uint32_t)lane < cnt;
        const uint32_t j = valid ? queue[lane] : 0u;
        const uint32_t insn = valid ? L[j] : 0u;
        const uint32_t op = insn_op(insn), rd = insn_rd(insn), rs = insn_rs(insn), rt = insn_rt(insn);
        const uint32_t imm = insn_imm(insn);
        const uint32_t other = pm & ~(1u << rd);
        const uint32_t sa = op == OP_ADDI ? rd : rs;
        bool rd_ok = valid;
        bool fault = false;
        bool pending = valid;
        for (;;) {
            const uint64_t pb = bal(pending);
            if (pb == 0ull) break;
            const uint32_t op0 = __builtin_amdgcn_readlane(op, (int)__ffsll((long long)pb) - 1);
            const bool mine = pending && op == op0;
            switch (op0) {
#define PCA_CASE(OPC) \
    case OPC: test_loop<OPC>(task, P, skipm, mine, rd, sa, rt, imm, other, rd_ok, fault, n_tests); break;
            PCA_CASE(OP_ADD)
            PCA_CASE(OP_SUB)
            PCA_CASE(OP_MUL)
            PCA_CASE(OP_DIV)
            PCA_CASE(OP_MOD)
            PCA_CASE(OP_AND)
            PCA_CASE(OP_OR)
            PCA_CASE(OP_XOR)
            PCA_CASE(OP_SHL)
            PCA_CASE(OP_SHR)
            PCA_CASE(OP_ASR)
            PCA_CASE(OP_LDI)
            PCA_CASE(OP_ADDI)
#undef PCA_CASE
            default: break;
            }
            pending = pending && !mine;
        }
        const uint32_t match = valid ? (other | (rd_ok ? (1u << rd) : 0u)) & 0xFEu : 0u;
        if (match) emit_hit(a, k, prefix_idx, jpath, j, match, (dead || fault) ? 1u : 0u);
    };

    uint32_t insn_next = lane < n ? L[lane] : 0u;
    for (int base = 0; base < n; base += 64) {
        const int j = base + lane;
        const bool valid = j < n;
        const uint32_t insn = insn_next;
        insn_next = j + 64 < n ? L[j + 64] : 0u;    // prefetch the next chunk
        const uint32_t op = insn_op(insn), rd = insn_rd(insn), rs = insn_rs(insn), rt = insn_rt(insn);
        const uint32_t imm = insn_imm(insn);
        const uint32_t la = live_after(live, insn);
        n_cand += valid ? (uint32_t)__popc(la & 0xFEu) : 0u;
        n_suf += valid;
        const bool is_cmp = op == OP_CMP;
        const uint32_t other = is_cmp ? pm : (pm & ~(1u << rd));
        bool alive;
        if (is_cmp || !valid) {
            // CMP writes no register: its hits are the parent's matches
            if (valid && other) emit_hit(a, k, prefix_idx, jpath, j, other & 0xFEu, dead);
            alive = false;
        } else if (two && other == 0u) {
            const uint32_t sa = op == OP_ADDI ? rd : rs;
            bool rd_ok = true;
            bool pending = true;
            for (;;) {
                const uint64_t pb = bal(pending);
                if (pb == 0ull) break;
                const uint32_t op0 = __builtin_amdgcn_readlane(op, (int)__ffsll((long long)pb) - 1);
                const bool mine = pending && op == op0;
                switch (op0) {
#define PCA_CASE(OPC) \
    case OPC: pair_eval<OPC>(task, P, skipm, mine, rd, sa, rt, imm, rd_ok); break;
                PCA_CASE(OP_ADD)
                PCA_CASE(OP_SUB)
                PCA_CASE(OP_MUL)
                PCA_CASE(OP_DIV)
                PCA_CASE(OP_MOD)
                PCA_CASE(OP_AND)
                PCA_CASE(OP_OR)
                PCA_CASE(OP_XOR)
                PCA_CASE(OP_SHL)
                PCA_CASE(OP_SHR)
                PCA_CASE(OP_ASR)
                PCA_CASE(OP_LDI)
                PCA_CASE(OP_ADDI)
#undef PCA_CASE
                default: break;
                }
                pending = pending && !mine;
            }
            n_tests += 2;
            alive = rd_ok;
        } else {
            alive = true;   // another register already matches (or T < 2): full test
        }
        // push survivors to the queue in lane order
        const uint64_t sb = bal(alive);
        if (alive) queue[qn + (uint32_t)__popcll(sb & ((1ull << lane) - 1ull))] = (uint16_t)j;
        qn += (uint32_t)__popcll(sb);
        if (qn >= 64) {
            wave_sync();
            drain(64);
            wave_sync();
            const uint32_t rest = qn - 64;
            if ((uint32_t)lane < rest) queue[lane] = queue[64 + lane];
            qn = rest;
            wave_sync();
        }
    }
    if (qn) {
        wave_sync();
        drain(qn);
        wave_sync();
    }
    const unsigned long long c0 = wave_sum(n_cand), c1 = wave_sum(n_suf), c2 = wave_sum(n_tests);
    if (lane == 0) {
        wc.c[0] += c0;
        wc.c[1] += c1;
        wc.c[2] += c2;
    }
}

// ---------------------------------------------- block-shared last position
// The children of one block (k = 2, 3) are the states of one node N after
// one more instruction each.  A child that writes one register x (an ALU,
// LDI or ADDI instruction) without branching or faulting differs from N
// only in x, on the tests it runs, and skips at the last position exactly
// the tests N would skip there (its resume and frozen bytes are N's).  A
// value group (op, a, b) with a != x and b != x therefore has the same
// value vector below the child as below N, the same rd values allowed on
// skipped tests (those rows are N's), and -- except for the rd = x bit of
// a DIV / MOD group whose divisor is zero on some test, which reads x's
// value -- the same outcome.  Those groups ("shared lanes") are decided
// once per block on N over every test; only the 15 lanes 8a + b with
// a == x or b == x are evaluated for each child, four children packed into
// one 64-lane pass (lane 16j + p: child j, p < 8 the pair (x, p), p >= 8
// the pair (p - 8, x)).  A CMP child writes no register: all its lanes are
// shared.  Children that branch, fault, or already hold a register equal
// to the target on every test take the per-child path.

// Per-wave buffers of the shared path, overlaid on pool slots 2 and 3 (a
// launch on this path cycles its pool through slots 0 and 1 only).
constexpr uint32_t GQ_CAP = 96;
constexpr uint32_t SH_KIDS = 32;   // shared-path children recorded per pass over a block
constexpr int SH_TA = 4;           // tests run on every group before the survivors queue (at most)
// One recorded child: its instruction and candidate index, the tests on
// which the register it writes equals the target, and that register's
// value on the first SH_TA tests.
struct ShareKid {
    uint32_t cc;              // instruction | candidate index at D-2 << 16
    uint32_t xm;              // bit t: x equals the target on test t
    uint16_t xv[SH_TA];
};
struct ShareLds {
    ShareKid kid[SH_KIDS];
    uint32_t gq[GQ_CAP];      // packed survivors: q | bit << 4 | ok << 10 | fault << 18 | child << 19
    uint32_t fn[64];          // per lane: N's groups matching every test (kind bits 0..12; 3, 4: DIV,
                              // MOD) | DIV rd mask << 13 | MOD rd mask << 20 (bits 1..7 of each)
                              // | DIV value match << 27 | MOD value match << 28
    uint32_t xs[8];           // [0]: N's group tests over all lanes; [x]: over the lanes reading x
    uint32_t zb[8];           // per register: tests (not skipped) on which N holds 0 there
};
static_assert(sizeof(ShareLds) <= 2 * MAXT * sizeof(TestState), "ShareLds fits in two pool slots");

// The value groups of one lane still matching: kind bits (ADD..ASR, ADDI,
// LDI; DIV and MOD kept as rd masks), and for a node scan whether DIV / MOD
// matched on every test whose divisor is nonzero (the rd = x bit of a
// child).
struct LaneKinds {
    uint32_t kl, okd, okm;
    bool fault, vmd, vmm;
};

// One test of every group of a lane (lbit = 8a + b: the LDI immediate index
// when below 16, b the ADDI immediate index); x, y the source values, tw
// the test's target word.  rmz() gives the register-match mask of the row
// this lane's groups leave alone on a zero divisor (called only when some
// lane divides by zero).  VM: keep the value-match flags of a node scan.
template <bool VM, class RMZ>
__device__ __forceinline__ void lane_test(LaneKinds &g, uint32_t x, uint32_t y, uint32_t tw, uint32_t lbit,
                                          uint32_t addi_imm, RMZ rmz)
{
    const uint32_t tg = tw & 0xFFFFu;
    const uint32_t sh = y & 15u;
    uint32_t keep = (lbit < 16u && ((tw >> (16u + lbit)) & 1u)) ? 1u << 12 : 0u;
    keep |= (((x + y) & 0xFFFFu) == tg ? 1u : 0u) << 0;
    keep |= (((x - y) & 0xFFFFu) == tg ? 1u : 0u) << 1;
    keep |= (((x * y) & 0xFFFFu) == tg ? 1u : 0u) << 2;
    keep |= ((x & y) == tg ? 1u : 0u) << 5;
    keep |= ((x | y) == tg ? 1u : 0u) << 6;
    keep |= ((x ^ y) == tg ? 1u : 0u) << 7;
    keep |= (((x << sh) & 0xFFFFu) == tg ? 1u : 0u) << 8;
    keep |= ((x >> sh) == tg ? 1u : 0u) << 9;
    keep |= (((uint32_t)((int32_t)(int16_t)(uint16_t)x >> sh) & 0xFFFFu) == tg ? 1u : 0u) << 10;
    keep |= (((x + addi_imm) & 0xFFFFu) == tg ? 1u : 0u) << 11;
    g.kl &= keep;
    if (any_lane((g.okd | g.okm) != 0u || (VM && (g.vmd || g.vmm)))) {
        const bool z = y == 0u;
        const uint32_t rmt = any_lane(z) ? rmz() : 0u;
        const int32_t sx = (int16_t)(uint16_t)x, sy = z ? 1 : (int32_t)(int16_t)(uint16_t)y;
        const int32_t qt = sdiv16_q(sx, sy);
        const bool dq = ((uint32_t)qt & 0xFFFFu) == tg, dr = ((uint32_t)(sx - qt * sy) & 0xFFFFu) == tg;
        g.okd &= z ? rmt : (dq ? 0xFFu : 0u);
        g.okm &= z ? rmt : (dr ? 0xFFu : 0u);
        g.fault = g.fault || (z && (g.okd != 0u || g.okm != 0u));
        if (VM) {
            g.vmd = g.vmd && (z || dq);
            g.vmm = g.vmm && (z || dr);
        }
    }
}

__device__ __forceinline__ uint32_t addi_imm_of(uint32_t b)
{
    return (b & 1u) ? (b >> 1) + 1u : 0x10000u - ((b >> 1) + 1u);
}

// Every value group of node N (state rows in LDS, live-mask index mN, tests
// skipped at the last position skipm) on every test, lanes = groups'
// source pairs.  Returns the group tests the per-child scan evaluates on
// this lane for one child (phase A up to TA tests while no skipped test
// has narrowed rd, then every test a group reaches); sets rm_and (AND of
// the skipped tests' register matches) and fn_any, and when some group
// matches every test writes the lanes' matches to sh.fn.
__device__ __forceinline__ uint32_t node_scan(const DevTask &task, const CandDev &cd, const TestState *N, uint32_t mN,
                                              uint64_t skipm, uint32_t act_tests, int TA, int lane, ShareLds &sh,
                                              uint32_t &rm_and_out, bool &fn_any, uint32_t &ran)
{
    const int T = task.T;
    const uint32_t ra = (uint32_t)lane >> 3, rb = (uint32_t)lane & 7u;
    // the tests (not skipped) where this lane's divisor rb is nonzero: the
    // DIV / MOD value-match flags change only there, so they keep the scan
    // going only while such a test is left (a divisor that is 0 on every
    // remaining test, r0 say, leaves them as they are)
    const uint32_t nzd = act_tests & ~(uint32_t)skipm & ~sh.zb[rb];
    const uint32_t addi_imm = addi_imm_of(rb);
    const uint32_t kb = cd.g->kind_bits[mN][lane];
    LaneKinds g;
    g.kl = kb & ~0x18u;
    g.okd = (kb >> 3 & 1u) ? 0xFEu : 0u;
    g.okm = (kb >> 4 & 1u) ? 0xFEu : 0u;
    g.vmd = (kb >> 3 & 1u) != 0u;
    g.vmm = (kb >> 4 & 1u) != 0u;
    g.fault = false;
    uint32_t rm_and = 0xFFu, ev = 0;
    int t = 0;
    for (; t < T; t++) {
        const bool vm_open = (g.vmd || g.vmm) && (nzd >> t) != 0u;
        if (!any_lane(g.kl != 0u || g.okd != 0u || g.okm != 0u || vm_open)) break;
        const uint32_t nal = (uint32_t)__popc(g.kl) + (g.okd != 0u ? 1u : 0u) + (g.okm != 0u ? 1u : 0u);
        if (t >= TA || rm_and == 0xFFu) ev += nal;
        ran += nal;
        const uint32_t tw = task.tgtw[t];
        if ((skipm >> t) & 1ull) {
            const uint32_t rmt = row_regmatch_uni(&N[t], tw & 0xFFFFu);
            g.okd &= rmt;
            g.okm &= rmt;
            rm_and &= rmt;
            if (rm_and == 0u) g.kl = 0u;
            continue;
        }
        const uint32_t x = N[t].r[ra], y = N[t].r[rb];
        lane_test<true>(g, x, y, tw, (uint32_t)lane, addi_imm,
                        [&]() { return row_regmatch_uni(&N[t], tw & 0xFFFFu); });
    }
    // the skipped tests after an early exit still narrow rd (a settled
    // DIV / MOD value match may still give its child a hit)
    for (; t < T; t++)
        if ((skipm >> t) & 1ull) rm_and &= row_regmatch_uni(&N[t], task.tgt[t]);
    rm_and_out = rm_and;
    const uint32_t fk = g.kl | ((g.okd != 0u || g.vmd) ? 1u << 3 : 0u) | ((g.okm != 0u || g.vmm) ? 1u << 4 : 0u);
    fn_any = any_lane(fk != 0u);
    if (fn_any)
        sh.fn[lane] = fk | ((g.okd & 0xFEu) << 12) | ((g.okm & 0xFEu) << 19) | (g.vmd ? 1u << 27 : 0u) |
                      (g.vmm ? 1u << 28 : 0u);
    return ev;
}

// Hits of one shared-path child on its shared lanes: N's groups matching
// every test, on the lanes that do not read x (x = 8: a CMP child, every
// lane).  xm: tests on which the child's x equals the target (bit t).
__device__ __forceinline__ void emit_shared(const ScanArgs &a, int k, const CandDev &cd, const ShareLds &sh,
                                            uint32_t x, uint32_t mP, uint32_t rm_and, uint32_t xm,
                                            uint32_t prefix_idx, const JPath &jp, uint32_t deadF, int lane)
{
    const uint32_t ra = (uint32_t)lane >> 3, rb = (uint32_t)lane & 7u;
    const bool isx = x < 8u && (ra == x || rb == x);
    const uint32_t f = isx ? 0u : sh.fn[lane];
    uint32_t fk = f & 0x1FFFu;
    if (!any_lane(fk != 0u)) return;
    const uint32_t zbl = sh.zb[rb];
    uint32_t okd = (f >> 12) & 0xFEu, okm = (f >> 19) & 0xFEu;
    if (x < 8u) {
        // rd = x of a DIV / MOD group: its value matched on every test with
        // a nonzero divisor, and x equals the target on the skipped tests
        // (rm_and) and on the tests where the divisor is zero
        const bool zx = (zbl & ~xm) == 0u && ((rm_and >> x) & 1u);
        okd = (okd & ~(1u << x)) | ((((f >> 27) & 1u) && zx) ? 1u << x : 0u);
        okm = (okm & ~(1u << x)) | ((((f >> 28) & 1u) && zx) ? 1u << x : 0u);
    }
    while (fk) {
        const uint32_t q = (uint32_t)__builtin_ctz(fk);
        fk &= fk - 1u;
        const bool dm = q == 3u || q == 4u;
        const uint32_t ok = q == 3u ? okd : (q == 4u ? okm : rm_and);
        if (!(ok & 0xFEu)) continue;
        const uint32_t g = cd.g->bits_gidx[mP][q][lane];
        if (g >= (uint32_t)VGMAX) continue;   // (no such group: cannot happen for a shared lane)
        const uint32_t okf = cd.g->vg_rdmask[mP][g] & ok;
        const uint32_t fl = deadF | ((dm && zbl != 0u) ? 1u : 0u);
        for (uint32_t r = 1; r < 8; r++)
            if (okf >> r & 1u) emit_hit(a, k, prefix_idx, jp, cd.g->vg_j[mP][g][r], 1u << r, fl);
    }
}

// The value the recorded child instruction c writes on test row `row`
// (the child runs there: not frozen, no pending branch past D-2).
__device__ __forceinline__ uint32_t kid_value(uint32_t c, const TestState &row)
{
    const uint32_t op = insn_op(c), rd = insn_rd(c);
    const uint32_t av = row.r[op == OP_ADDI ? rd : insn_rs(c)], bv = row.r[insn_rt(c)];
    uint32_t f;
    bool fault;
    return alu_eval(op, av, bv, insn_imm(c), row.flags, &f, &fault) & 0xFFFFu;
}

// Remaining tests (t0..T-1) of packed survivors gq[0..cnt), one per lane,
// against their child's state (N with x replaced by the child's value),
// with an early exit; survivors give one hit per allowed rd.  kid: the
// pass's first recorded child; lvp: N's live mask; cpos = D - 2.
__device__ __forceinline__ void group_drain(const DevTask &task, const ScanArgs &a, int k, const CandDev &cd,
                                            const TestState *N, const ShareKid *kid, uint32_t lvp, uint32_t cpos,
                                            const uint32_t *gq, uint64_t skipm, uint32_t cnt, int t0, int lane,
                                            WaveCounters &wc, uint32_t (*tst)[2], Pool &pool, uint32_t prefix_idx,
                                            uint32_t js, uint32_t deadF)
{
    const int T = task.T;
    const bool valid = (uint32_t)lane < cnt;
    const uint32_t e = valid ? gq[lane] : 0u;
    const uint32_t q = e & 15u, bit = (e >> 4) & 63u, jj = (e >> 19) & 3u;
    uint32_t ok = valid ? (e >> 10) & 0xFFu : 0u;
    bool flt = (e >> 18) & 1u;
    const ShareKid &kd = kid[jj];
    const uint32_t cc = kd.cc;
    const uint32_t ckid = cc & 0xFFFFu, cj = cc >> 16, xj = insn_rd(ckid);
    const uint32_t mj = ((lvp | (1u << xj)) >> 1) & 127u;
    const uint32_t xa = bit >> 3, yb = bit & 7u;
    const uint32_t aimm = addi_imm_of(yb);
    const uint32_t aimm2 = (aimm & 0xFFFFu) * 0x10001u;
    const bool divmod = q == OP_DIV - OP_ADD || q == OP_MOD - OP_ADD;
    bool any_div = any_lane(divmod && ok != 0u);
    const bool tst8 = tst && (pool.nd & 7u) == 0u;
    pool.nd++;
    uint32_t n_ev = 0;
    // the child's x on test t: recorded for the first SH_TA tests, else
    // computed from N's row
    auto xval = [&](int t) -> uint32_t {
        const TestState &row = N[t];
        if (t < SH_TA) return kd.xv[t];
        const bool runs = row.frozen == 0u && (uint32_t)row.resume <= cpos;
        return runs ? kid_value(ckid, row) : (uint32_t)row.r[xj];
    };
    auto val = [&](uint32_t r, int t) -> uint32_t { return r == xj ? xval(t) : (uint32_t)N[t].r[r]; };
    for (int t_ = t0; t_ < T; t_++) {
        const int tt = (int)uni((uint32_t)t_);
        const uint64_t al = bal(ok != 0u);
        if (al == 0ull) break;
        const uint32_t before = (uint32_t)__popcll(al);
        n_ev += before;
        if (any_div) any_div = any_lane(divmod && ok != 0u);
        if (!any_div && tt + 1 < T && ((skipm >> tt) & 3ull) == 0ull) {
            const uint32_t tw0 = task.tgtw[tt], tw1 = task.tgtw[tt + 1];
            const uint32_t tg2 = (tw0 & 0xFFFFu) | (tw1 << 16);
            const uint32_t x2 = val(xa, tt) | (val(xa, tt + 1) << 16);
            const uint32_t y2 = val(yb, tt) | (val(yb, tt + 1) << 16);
            const u16x2 X = __builtin_bit_cast(u16x2, x2), Y = __builtin_bit_cast(u16x2, y2);
            const u16x2 S = Y & (unsigned short)15;
            uint32_t v = __builtin_bit_cast(uint32_t, X + Y);
            v = q == 1u ? __builtin_bit_cast(uint32_t, X - Y) : v;
            v = q == 2u ? __builtin_bit_cast(uint32_t, X * Y) : v;
            v = q == 5u ? (x2 & y2) : v;
            v = q == 6u ? (x2 | y2) : v;
            v = q == 7u ? (x2 ^ y2) : v;
            v = q == 8u ? __builtin_bit_cast(uint32_t, X << S) : v;
            v = q == 9u ? __builtin_bit_cast(uint32_t, X >> S) : v;
            v = q == 10u ? __builtin_bit_cast(uint32_t, __builtin_bit_cast(i16x2, x2) >> __builtin_bit_cast(i16x2, S))
                         : v;
            v = q == 11u ? __builtin_bit_cast(uint32_t, X + __builtin_bit_cast(u16x2, aimm2)) : v;
            const uint32_t d = v ^ tg2;
            bool m0 = (d & 0xFFFFu) == 0u, m1 = (d >> 16) == 0u;
            if (q == 12u) {
                m0 = (tw0 >> (16 + bit)) & 1u;
                m1 = (tw1 >> (16 + bit)) & 1u;
            }
            const uint32_t mid = (uint32_t)__popcll(bal(ok != 0u && m0));
            n_ev += mid;
            if (!(m0 && m1)) ok = 0;
            if (tst8) {
                const uint32_t after = (uint32_t)__popcll(bal(ok != 0u));
                if (lane == 0) {
                    tst[tt][0] += before;
                    tst[tt][1] += before - mid;
                    tst[tt + 1][0] += mid;
                    tst[tt + 1][1] += mid - after;
                }
            }
            t_ = tt + 1;
            continue;
        }
        const bool sk = (skipm >> tt) & 1ull;
        const uint32_t tw = task.tgtw[tt];
        const uint32_t tg = tw & 0xFFFFu;
        if (sk) {
            // the child did not run this test: its row is N's
            ok &= row_regmatch_uni(&N[tt], tg);
        } else {
            const uint32_t x = val(xa, tt), y = val(yb, tt);
            const uint32_t sh = y & 15u;
            uint32_t v = x + y;
            v = q == 1u ? x - y : v;
            v = q == 2u ? x * y : v;
            v = q == 5u ? (x & y) : v;
            v = q == 6u ? (x | y) : v;
            v = q == 7u ? (x ^ y) : v;
            v = q == 8u ? x << sh : v;
            v = q == 9u ? x >> sh : v;
            v = q == 10u ? (uint32_t)((int32_t)(int16_t)(uint16_t)x >> sh) : v;
            v = q == 11u ? x + aimm : v;
            bool match = (v & 0xFFFFu) == tg;
            if (q == 12u) match = (tw >> (16 + bit)) & 1u;
            if (any_div && divmod) {
                if (y == 0u) {
                    // a zero divisor stops the VM: rd keeps the child's value
                    const uint32_t rn = row_regmatch(&N[tt], tg);
                    ok &= (rn & ~(1u << xj)) | ((xval(tt) == tg ? 1u : 0u) << xj);
                    flt = true;
                    match = true;
                } else {
                    const int32_t sx = (int16_t)(uint16_t)x, sy = (int16_t)(uint16_t)y;
                    const int32_t qt = sdiv16_q(sx, sy);
                    match = ((uint32_t)(q == OP_DIV - OP_ADD ? qt : sx - qt * sy) & 0xFFFFu) == tg;
                }
            }
            if (!match) ok = 0;
        }
        if (tst8) {
            const uint32_t after = (uint32_t)__popcll(bal(ok != 0u));
            if (lane == 0) {
                tst[tt][0] += before;
                tst[tt][1] += before - after;
            }
        }
    }
    if (lane == 0) wc.c[2] += n_ev;
    if (ok != 0u) {
        const uint32_t g = cd.g->bits_gidx[mj][q][bit];
        const uint32_t okf = g < (uint32_t)VGMAX ? cd.g->vg_rdmask[mj][g] & ok : 0u;
        const uint32_t fl = deadF | ((divmod && flt) ? 1u : 0u);
        JPath jp;
        jp.lds = nullptr;
        jp.j0 = k == 3 ? js : cj;
        jp.j1 = cj;
        for (uint32_t r = 1; r < 8; r++)
            if (okf >> r & 1u) emit_hit(a, k, prefix_idx, jp, cd.g->vg_j[mj][g][r], 1u << r, fl);
    }
}

// The lanes that read x of gn (<= 4) recorded children kid[0..gn): phase A
// over the first TA (<= SH_TA) tests, then the survivors through
// group_drain.
__device__ __forceinline__ void run_group(const DevTask &task, const ScanArgs &a, int k, const CandDev &cd,
                                          const TestState *N, ShareLds &sh, WaveLds &sv, const ShareKid *kid,
                                          uint32_t gn, uint32_t lvp, uint32_t cpos, uint64_t skipm, int TA, int lane,
                                          WaveCounters &wc, uint32_t (*tst)[2], Pool &pool, uint32_t prefix_idx,
                                          uint32_t js, uint32_t deadF)
{
    const uint32_t j = (uint32_t)lane >> 4, p = (uint32_t)lane & 15u;
    const bool in = j < gn;
    const ShareKid &kd = kid[in ? j : 0u];
    const uint32_t xj = in ? insn_rd(kd.cc) : 0u;
    const uint32_t mj = ((lvp | (1u << xj)) >> 1) & 127u;
    const bool lo = p < 8u;
    const uint32_t o = p & 7u;
    // (x, x) appears at p = x and p = 8 + x: the second copy is idle
    const bool valid = in && (lo || o != xj);
    const uint32_t ra = lo ? xj : o, rb = lo ? o : xj;
    const uint32_t bit = (ra << 3) | rb;
    const uint32_t kb = valid ? cd.g->kind_bits[mj][bit] : 0u;
    LaneKinds g;
    g.kl = kb & ~0x18u;
    g.okd = (kb >> 3 & 1u) ? 0xFEu : 0u;
    g.okm = (kb >> 4 & 1u) ? 0xFEu : 0u;
    g.fault = g.vmd = g.vmm = false;
    const uint32_t addi_imm = addi_imm_of(rb);
    uint32_t rm_and = 0xFFu, ev = 0;
    for (int t = 0; t < TA; t++) {
        const uint32_t nal = (uint32_t)__popc(g.kl) + (g.okd != 0u ? 1u : 0u) + (g.okm != 0u ? 1u : 0u);
        if (!any_lane(nal != 0u)) break;
        if (rm_and == 0xFFu) ev += nal;
        const uint32_t tw = task.tgtw[t];
        if ((skipm >> t) & 1ull) {
            const uint32_t rmt = row_regmatch_uni(&N[t], tw & 0xFFFFu);
            g.okd &= rmt;
            g.okm &= rmt;
            rm_and &= rmt;
            continue;
        }
        const uint32_t xv = kd.xv[t];
        const uint32_t nv = N[t].r[o];
        const uint32_t ov = o == xj ? xv : nv;
        const uint32_t X = lo ? xv : ov, Y = lo ? ov : xv;
        lane_test<false>(g, X, Y, tw, bit, addi_imm, [&]() {
            const uint32_t tg = tw & 0xFFFFu;
            const uint32_t rn = row_regmatch_uni(&N[t], tg);
            return (rn & ~(1u << xj)) | ((xv == tg ? 1u : 0u) << xj);
        });
        if (tst && t == 2) {
            const uint32_t left = (uint32_t)__popc(g.kl) + (g.okd != 0u ? 1u : 0u) + (g.okm != 0u ? 1u : 0u);
            __hip_atomic_fetch_add(&sv.pa2[lane], (unsigned long long)nal | ((unsigned long long)(nal - left) << 32),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
        }
    }
    if (ev) __hip_atomic_fetch_add(&sv.evl[lane], ev, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    if (rm_and == 0u) return;
    const uint32_t sm = g.kl | (g.okd != 0u ? 1u << 3 : 0u) | (g.okm != 0u ? 1u << 4 : 0u);
    const uint32_t c = (uint32_t)__popc(sm);
    const uint32_t incl = wave_incl_scan32(c);
    const uint32_t total = uni(__builtin_amdgcn_readlane(incl, 63));
    if (total == 0u) return;
    const uint32_t excl = incl - c;
    uint32_t qn = 0, pushed = 0;
    for (;;) {
        if (pushed < total) {
            const uint32_t take = min(GQ_CAP - qn, total - pushed);
            if (c != 0u && excl < pushed + take && excl + c > pushed) {
                uint32_t mk = sm, gi = excl;
                while (mk) {
                    const uint32_t q = (uint32_t)__builtin_ctz(mk);
                    mk &= mk - 1u;
                    if (gi >= pushed && gi < pushed + take) {
                        const uint32_t okq = q == 3u ? g.okd : (q == 4u ? g.okm : rm_and);
                        sh.gq[qn + (gi - pushed)] = q | (bit << 4) | (okq << 10) |
                                                    (((q == 3u || q == 4u) && g.fault) ? 1u << 18 : 0u) | (j << 19);
                    }
                    gi++;
                }
            }
            qn += take;
            pushed += take;
        }
        const bool all = pushed == total;
        const uint32_t cnt = qn >= 64u ? 64u : (all ? qn : 0u);
        if (cnt) {
            wave_sync();
            group_drain(task, a, k, cd, N, kid, lvp, cpos, sh.gq, skipm, cnt, TA, lane, wc, tst, pool, prefix_idx, js,
                        deadF);
            wave_sync();
            const uint32_t rest = qn - cnt;
            if ((uint32_t)lane < rest) sh.gq[lane] = sh.gq[64 + lane];
            qn = rest;
            wave_sync();
        }
        if (all && qn == 0u) break;
    }
}

// k == 0: the prefix is the whole kernel; check its registers directly.
__device__ __forceinline__ void leaf_parent(const DevTask &task, const ScanArgs &a, const TestState *P, uint32_t live,
                            uint32_t prefix_idx, int lane, WaveCounters &wc)
{
    const int T = task.T;
    const bool act = lane < T;
    TestState st;
    if (act) st = P[lane];
    const uint32_t tgt = act ? task.tgt[lane] : 0u;
    uint32_t pm = 0;
    for (int r = 1; r < 8; r++) pm |= (bal(act && st.r[r] != tgt) == 0ull ? 1u : 0u) << r;
    const bool cyc = act && ((uint32_t)st.exec + 2u > (uint32_t)task.max_cycles);
    const uint32_t dead = bal(act && (st.frozen || cyc)) != 0ull ? 1u : 0u;
    if (lane == 0) {
        wc.c[3]++;
        wc.c[0] += (unsigned)__popc(live & 0xFEu);
        wc.c[1]++;
        if (pm & 0xFEu) {
            const uint32_t slot = atomicAdd(a.hit_count, 1u);
            if (slot < a.hit_cap) {
                HitRec h;
                h.prefix = prefix_idx;
                h.mask = (uint8_t)(pm & 0xFEu);
                h.flags = (uint8_t)dead;
                for (int q = 0; q < PCAGPU_MAX_KDEPTH; q++) h.j[q] = 0;
                h.pad = 0;
                a.hits[slot] = h;
            }
        }
    }
}

// MULTI = false: one item and its task record, passed by value in the kernel
// arguments (the single-task path; no per-unit item lookup, the context is
// read from the kernel-argument segment).  MULTI = true: lch.items[0..n_items)
// in device memory, the item of each unit found by binary search over
// unit_begin.  Per wave, dynamic LDS holds the state levels above the
// scanned parents (max(lch.levels, 2) - 1 of them); the scanned parents
// themselves (level k - 1) sit in the NSLOT slots of the wave's WaveLds,
// where they stay while their survivors wait in the pooled queue.
template <bool MULTI, int KF, bool SHARED>
__device__ __forceinline__ void scan_body(const ScanLaunch &lch, const ScanArgs &a1, const DevTask &task1)
{
    // KF: the suffix length of every item, fixed at compile time (2 or 3;
    // 0: read from the item).  SHARED: blocks of 2- and 3-instruction
    // suffixes scan the last position shared per block (items whose tile
    // asks for it; the pool then uses slots 0 and 1)
    constexpr bool K23 = KF != 0;
    extern __shared__ TestState s_dyn[];
    __shared__ WaveLds s_wave[SCAN_WAVES];
    __shared__ int s_cnt[128], s_split[128];
    __shared__ uint32_t s_ctot[128];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WaveLds &sv = s_wave[w];
    for (int i = lane; i < MAXT * 2; i += 64) (&sv.tst[0][0])[i] = 0u;
    if (lane < 4) sv.wc[lane] = 0ull;
    if (lane < 2) sv.wx[lane] = 0ull;
    sv.evl[lane] = 0u;
    sv.pa2[lane] = 0ull;
    uint32_t (*tst)[2] = (!MULTI && lch.tstat) ? sv.tst : nullptr;
    // per-mask tables live in LDS: no global round trip per parent
    for (int i = threadIdx.x; i < 128; i += SCAN_WG) {
        s_cnt[i] = lch.cand.count[i];
        s_split[i] = lch.cand.split[i];
        s_ctot[i] = lch.cand.cand_total[i];
    }
    __syncthreads();
    CandDev cand = lch.cand;
    cand.count = s_cnt;
    cand.split = s_split;
    cand.cand_total = s_ctot;
    const uint32_t nreg = (lch.levels > 2u ? lch.levels : 2u) - 1u;
    auto lev = [&](int L) -> TestState * { return s_dyn + ((uint32_t)w * nreg + (uint32_t)L) * MAXT; };
    Pool pool;
    pool.qn = pool.pend = pool.cur = pool.nd = 0;
    pool.ns = SHARED ? 2u : NSLOT;

    WaveCounters wc;
    wc.c = sv.wc;
    // current item (wave-uniform) and its context: the one-item kernel reads
    // its work item from the kernel arguments (a1), where the fields stay
    // (reloaded through the scalar cache, not held in registers); the
    // multi-item kernel copies the current item's fields into am
    uint64_t it_begin = 0, it_end = 0;
    const DevTask *taskp = nullptr;
    ScanArgs am;
    am.counters = lch.counters;
    am.work = lch.work;
    const ScanArgs &a = MULTI ? am : a1;
    uint32_t cur_i = 0xFFFFFFFFu, cur_js = 0xFFFFFFFFu;   // cached levels 0 / 1
    uint32_t my_tgt = 0;                                   // lane = test: its target
    if constexpr (!MULTI) {
        taskp = &task1;
        my_tgt = lane < task1.T ? task1.tgt[lane] : 0u;
        it_begin = 0;
        it_end = a1.n_units;
    }
    const uint64_t grab = lch.grab > 1u ? lch.grab : 1u;
    uint64_t ug_next = 0, ug_end = 0;   // the wave's current run of units
    // the pooled survivors of the item the wave leaves (fewer than 64); the
    // one-item kernel has a single such drain, after its unit loop
    auto flush_pool = [&]() {
        const int ta = taskp->T < (int)a.tile.mask_tests ? taskp->T : (int)a.tile.mask_tests;
        wave_sync();
        pool_drain(*taskp, a, KF ? KF : a.k, cand, sv, pool.qn, ta, lane, wc, tst, pool);
        wave_sync();
        pool.qn = 0;
    };
    for (;;) {
        if (ug_next >= ug_end) {
            unsigned long long g0 = 0;
            if (lane == 0) g0 = atomicAdd(lch.work, (unsigned long long)grab);
            g0 = uni64(g0);
            if (g0 >= lch.n_units) break;
            ug_next = g0;
            ug_end = min((uint64_t)g0 + grab, (uint64_t)lch.n_units);
        }
        const uint64_t ug = ug_next++;
        if (MULTI && (ug < it_begin || ug >= it_end)) {
            if constexpr (MULTI) {
                if (pool.qn) flush_pool();
            }
            // the item holding this unit: largest ii with unit_begin <= ug
            uint32_t lo = 0, hi = lch.n_items;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (lch.items[mid].unit_begin <= ug) lo = mid; else hi = mid;
            }
            const ScanItem &it = lch.items[uni(lo)];
            taskp = it.task;
            am.prefixes = it.prefixes;
            am.n_prefix = it.n_prefix;
            am.blocks = it.blocks;
            am.n_units = it.n_units;
            am.D = it.D;
            am.tile = it.tile;
            am.s = it.s;
            am.k = it.k;
            am.hits = it.hits;
            am.hit_cap = it.hit_cap;
            am.hit_count = it.hit_count;
            it_begin = it.unit_begin;
            it_end = it.unit_begin + it.n_units;
            cur_i = 0xFFFFFFFFu;
            cur_js = 0xFFFFFFFFu;
            my_tgt = lane < taskp->T ? taskp->tgt[lane] : 0u;
        }
        const DevTask &task = *taskp;
        const int k = KF ? KF : a.k;
        const uint64_t u = ug - it_begin;
        uint32_t pi, js = 0, j0 = 0;
        if (k >= 2) {
            const ScanBlock B = a.blocks[u];
            pi = uni(B.prefix);
            js = uni(B.js);
            j0 = uni(B.j0);
        } else {
            pi = uni((uint32_t)u);
        }
        const pca_prefix_t *P = &a.prefixes[pi];
        const uint32_t live0 = uni(P->live);
        // the scanned parents (level k - 1) sit in the slots: for k == 1 the
        // prefix itself, replayed into the current slot every time
        auto lvl = [&](int L) -> TestState * { return L == k - 1 ? sv.slot[pool.cur] : lev(L); };
        if ((!K23 && k == 1) || pi != cur_i) {
            replay_prefix(task, P, a.s, (!K23 && k == 1) ? lvl(0) : lev(0), lane);
            cur_i = k == 1 ? 0xFFFFFFFFu : pi;
            cur_js = 0xFFFFFFFFu;
            wave_sync();
        }
        if constexpr (!K23) {
            if (k == 0) {
                leaf_parent(task, a, lev(0), live0, pi, lane, wc);
                continue;
            }
        }
        if (lane == 0) sv.live[0] = live0;
        // level of the scanned parents' first level: 0 (the prefix itself,
        // k == 1), 1 (k == 2), 2 (k >= 3); the block's children sit there
        const int Lu = k <= 1 ? 0 : (k == 2 ? 1 : 2);
        uint32_t nb = 1, lvp = live0, cw = 0;
        if (k >= 3 && js != cur_js) {
            const uint32_t c1 = uni(cand_at(cand, live0, a.D - a.s - 1, (int)js));
            step_level(task, lev(0), lev(1), c1, a.s, lane);
            if (lane == 0) sv.live[1] = live_after(live0, c1);
            cur_js = js;
        }
        if (k >= 2) {
            if (k >= 3) {
                wave_sync();
                lvp = uni(sv.live[1]);
                if (lane == 0) sv.idx[1] = js;
            }
            const int mfp = a.D - a.s - Lu;   // children at position s + Lu - 1
            const int nc = cand_count(cand, lvp, mfp);
            nb = (int)j0 < nc ? (uint32_t)min(nc - (int)j0, (int)a.tile.block_kids) : 0u;
            // the block's candidate words, one per lane, in one coalesced read
            cw = (uint32_t)lane < nb ? cand_at(cand, lvp, mfp, (int)j0 + lane) : 0u;
        }
        if (K23 || (MULTI && (k == 2 || k == 3))) {
            // every child of the block is a scanned parent (kernel position
            // D-2): stepped from the block's parent state, held in registers,
            // into the current slot (without flags: nothing at the last
            // position reads them), summarised from registers, scanned
            const bool act = lane < task.T;
            PackedState pl = {};
            if (act) pl = load_ps(&lev(Lu - 1)[lane]);
            const uint32_t cpos = (uint32_t)(a.D - 2);
            if (SHARED && a.tile.mode == TILE_SHARED && nb != 0u) {
                // the block's node N (level Lu - 1, in LDS and in pl)
                const TestState *N = lev(Lu - 1);
                ShareLds &sh = *reinterpret_cast<ShareLds *>(&sv.slot[2][0]);
                const uint32_t lpos = (uint32_t)(a.D - 1);
                const int TA0 = task.T < (int)a.tile.mask_tests ? task.T : (int)a.tile.mask_tests;
                const int TA = TA0 < SH_TA ? TA0 : SH_TA;
                const uint64_t am = bal(act);
                const uint32_t w4 = pl.w[4];
                // what every shared-path child inherits from N: the tests
                // it skips at the last position, its frozen tests, and
                // whether it is dead (frozen, or out of cycles once it has
                // run its own instruction)
                const uint64_t skipN = bal(((w4 >> 8) & 0xFFFFu) > lpos) & am;
                const uint64_t frzN = bal((w4 & 0xFF0000u) != 0u) & am;
                const uint32_t exP = ((w4 >> 24) + ((act && ps_runs(w4, cpos)) ? 1u : 0u)) & 0xFFu;
                const uint32_t mc = (uint32_t)task.max_cycles;
                const uint64_t cycF = (bal(exP + 2u > mc) | (bal(exP + 3u > mc) & ~skipN)) & am;
                const uint32_t deadF = (frzN | cycF) != 0ull ? 1u : 0u;
                const uint32_t pmN = summarize_state(task, pl, act, lpos, my_tgt).pm;
                uint32_t rm_andN = 0xFFu, ran = 0;
                bool fn_any = false;
                {
                    // per register, the tests (not skipped) where N holds 0
                    uint32_t zr = 0;
#pragma unroll
                    for (uint32_t r = 0; r < 8; r++) {
                        const uint64_t zm = bal(ps_reg(pl, r) == 0u) & am & ~skipN;
                        zr = lane == (int)r ? (uint32_t)zm : zr;
                    }
                    if (lane < 8) sh.zb[lane] = zr;
                    wave_sync();
                }
                const uint32_t evn = node_scan(task, cand, N, (lvp >> 1) & 127u, skipN, (uint32_t)am, TA, lane, sh,
                                               rm_andN, fn_any, ran);
                // N's group tests over all lanes and over the lanes reading
                // each register (a child is credited the lanes it shares)
                const uint32_t ev_all = (uint32_t)wave_sum((unsigned long long)evn);
                const unsigned long long ran_all = wave_sum((unsigned long long)ran);
                if (lane == 0) sv.wx[1] += ran_all;
                wave_sync();
                sh.gq[lane] = evn;
                wave_sync();
                if (lane < 8) {
                    uint32_t xsum = 0;
                    for (int i = 0; i < 8; i++) xsum += sh.gq[lane * 8 + i] + (i != lane ? sh.gq[i * 8 + lane] : 0u);
                    sh.xs[lane] = lane == 0 ? ev_all : xsum;
                }
                wave_sync();
                const uint32_t jsb = k == 3 ? js : 0u;
                const uint32_t mN = (lvp >> 1) & 127u;
                // a shared block holds up to 128 children: the second 64
                // candidate words, one per lane (kid_word picks by index)
                uint32_t cw2 = 0;
                if (nb > 64u) {
                    const int mfp2 = a.D - a.s - (k == 2 ? 1 : 2);
                    cw2 = (uint32_t)lane + 64u < nb ? cand_at(cand, lvp, mfp2, (int)j0 + 64 + lane) : 0u;
                }
                auto kid_word = [&](uint32_t ci) -> uint32_t {
                    return ci < 64u ? __builtin_amdgcn_readlane(cw, (int)ci)
                                    : __builtin_amdgcn_readlane(cw2, (int)(ci - 64u));
                };
                // the tests on which the children run (not frozen, no
                // pending branch past D-2)
                const bool runs2 = act && ps_runs(w4, cpos);
                const uint64_t runs2m = bal(runs2);
                for (uint32_t base = 0; base < nb; base += SH_KIDS) {
                    const uint32_t nh = min(nb - base, SH_KIDS);
                    // pass 1: sort the children; record the ones that
                    // write a register (their value from N's row: one ALU
                    // operation on the uniform opcode, no general step),
                    // count the shared ones per register written (hist: 7
                    // bits per register, slot 0 for CMP), defer the
                    // per-child ones (bit ci - base of gmask)
                    uint32_t nx = 0, gmask = 0, cmask = 0;
                    uint64_t hist = 0;
                    for (uint32_t ci = base; ci < base + nh; ci++) {
                        const uint32_t c = kid_word(ci);
                        const uint32_t op = insn_op(c), x = insn_rd(c);
                        const uint32_t cbit = 1u << (ci - base);
                        if (op == OP_BR) {
                            gmask |= cbit;
                            continue;
                        }
                        if (op == OP_CMP) {
                            // writes no register: every lane shared
                            if (pmN & 0xFEu) {
                                gmask |= cbit;
                            } else {
                                cmask |= cbit;
                                hist += 1ull;
                            }
                            continue;
                        }
                        const TestState &row = N[act ? lane : 0];
                        const uint32_t av = row.r[op == OP_ADDI ? x : insn_rs(c)], bv = row.r[insn_rt(c)];
                        const uint32_t ov = row.r[x];
                        // a DIV / MOD by zero on a test it runs freezes it
                        if ((op == OP_DIV || op == OP_MOD) && (bal(bv == 0u) & runs2m) != 0ull) {
                            gmask |= cbit;
                            continue;
                        }
                        const uint32_t imm = insn_imm(c);
                        uint32_t v = ov;
                        switch (op) {
#define PCA_KV(OPC) \
    case OPC: v = op_value<OPC>(av, bv, imm); break;
                        PCA_KV(OP_ADD)
                        PCA_KV(OP_SUB)
                        PCA_KV(OP_MUL)
                        PCA_KV(OP_DIV)
                        PCA_KV(OP_MOD)
                        PCA_KV(OP_AND)
                        PCA_KV(OP_OR)
                        PCA_KV(OP_XOR)
                        PCA_KV(OP_SHL)
                        PCA_KV(OP_SHR)
                        PCA_KV(OP_ASR)
                        PCA_KV(OP_LDI)
                        PCA_KV(OP_ADDI)
#undef PCA_KV
                        default: break;
                        }
                        const uint32_t xv = runs2 ? v : ov;
                        const uint64_t hitm = bal(xv == my_tgt) & am;
                        if (((pmN & ~(1u << x)) | (hitm == am ? 1u << x : 0u)) & 0xFEu) {
                            gmask |= cbit;
                            continue;
                        }
                        hist += 1ull << (7u * x);
                        ShareKid &kd = sh.kid[nx];
                        if (lane == 0) {
                            kd.cc = c | ((ci + j0) << 16);
                            kd.xm = (uint32_t)hitm;
                        }
                        if (lane < SH_TA) kd.xv[lane] = (uint16_t)xv;
                        nx++;
                    }
                    {
                        // counters of the shared children: per register x
                        // written (lane x; lane 0: CMP) their count times
                        // the candidates, suffixes and shared group tests
                        // of one such child
                        const uint32_t xl = (uint32_t)lane & 7u;
                        const uint32_t cnt = lane < 8 ? (uint32_t)(hist >> (7u * xl)) & 127u : 0u;
                        const uint32_t mx = xl == 0u ? mN : ((lvp | (1u << xl)) >> 1) & 127u;
                        const uint32_t credit = sh.xs[0] - (xl == 0u ? 0u : sh.xs[xl]);
                        const unsigned long long c0 = wave_sum((unsigned long long)(cnt * cand.cand_total[mx]));
                        const unsigned long long c1 = wave_sum((unsigned long long)(cnt * (uint32_t)cand.count[mx]));
                        const unsigned long long c2 = wave_sum((unsigned long long)cnt * credit);
                        if (lane == 0) {
                            wc.c[0] += c0;
                            wc.c[1] += c1;
                            wc.c[2] += c2;
                            sv.wx[0] += c2;
                        }
                    }
                    wave_sync();
                    // pass 2: the shared lanes' hits (only when some group
                    // of N matches every test)
                    if (fn_any) {
                        for (uint32_t i = 0; i < nx; i++) {
                            const uint32_t cc = sh.kid[i].cc, xm = sh.kid[i].xm;
                            const uint32_t x = insn_rd(cc & 0xFFFFu);
                            JPath jp;
                            jp.lds = nullptr;
                            jp.j0 = k == 3 ? js : cc >> 16;
                            jp.j1 = cc >> 16;
                            emit_shared(a, k, cand, sh, x, ((lvp | (1u << x)) >> 1) & 127u, rm_andN, xm, pi, jp,
                                        deadF, lane);
                        }
                        for (uint32_t mk = cmask; mk; mk &= mk - 1u) {
                            const uint32_t ci = base + (uint32_t)__builtin_ctz(mk);
                            JPath jp;
                            jp.lds = nullptr;
                            jp.j0 = k == 3 ? js : j0 + ci;
                            jp.j1 = j0 + ci;
                            emit_shared(a, k, cand, sh, 8u, mN, rm_andN, 0u, pi, jp, deadF, lane);
                        }
                    }
                    // pass 3: the lanes that read x, four children a pass
                    for (uint32_t g0 = 0; g0 < nx; g0 += 4u) {
                        run_group(task, a, k, cand, N, sh, sv, &sh.kid[g0], min(nx - g0, 4u), lvp, cpos, skipN, TA,
                                  lane, wc, tst, pool, pi, jsb, deadF);
                        wave_sync();
                    }
                    // pass 4: the per-child path (branches, faults, a
                    // register already equal to the target)
                    for (uint32_t mk = gmask; mk; mk &= mk - 1u) {
                        const uint32_t ci = base + (uint32_t)__builtin_ctz(mk);
                        const uint32_t c = kid_word(ci);
                        PackedState st = pl;
                        if (act && ps_runs(st.w[4], cpos)) step_packed<false, false>(st, c, cpos);
                        TestState *PP = sv.slot[pool.cur];
                        if (act) store_ps(&PP[lane], st);
                        const ParentSummary ps = summarize_state(task, st, act, lpos, my_tgt);
                        const uint32_t lvL = live_after(lvp, c);
                        JPath jpath;
                        jpath.lds = nullptr;
                        jpath.j0 = k == 3 ? js : j0 + ci;
                        jpath.j1 = j0 + ci;
                        wave_sync();
                        if (ps.pm == 0u)
                            scan_parent_bits(task, a, k, cand, PP, lvL, pi, jpath, lane, wc, sv, pool, tst, ps);
                        else
                            scan_parent(task, a, k, cand, PP, lvL, pi, jpath, lane, wc, sv.cq, ps);
                        wave_sync();
                    }
                }
                if (lane == 0) wc.c[3] += nb;
                continue;
            }
            for (uint32_t ci = 0; ci < nb; ci++) {
                const uint32_t c = __builtin_amdgcn_readlane(cw, (int)ci);
                TestState *PP = sv.slot[pool.cur];
                PackedState st = pl;
                // the parent's status word enters each child as an unknown
                // value: the tests on its flags, resume and frozen bytes are
                // made inside the branch of the child's opcode that reads
                // them, instead of hoisted out of the loop as lane masks
                // held (and spilled) across it
                __asm__ volatile("" : "+v"(st.w[4]));
                // one region for the lanes whose test runs the child (the
                // others keep the parent's words): no per-condition copies
                if (act && ps_runs(st.w[4], cpos)) step_packed<false, false>(st, c, cpos);
                if (act) store_ps(&PP[lane], st);
                const ParentSummary ps = summarize_state(task, st, act, (uint32_t)(a.D - 1), my_tgt);
                const uint32_t lvL = live_after(lvp, c);
                JPath jpath;
                jpath.lds = nullptr;
                jpath.j0 = k == 3 ? js : j0 + ci;
                jpath.j1 = j0 + ci;
                wave_sync();
                if (ps.pm == 0u)
                    scan_parent_bits(task, a, k, cand, PP, lvL, pi, jpath, lane, wc, sv, pool, tst, ps);
                else
                    scan_parent(task, a, k, cand, PP, lvL, pi, jpath, lane, wc, sv.cq, ps);
                wave_sync();
            }
            if (lane == 0) wc.c[3] += nb;
            continue;
        }
        if constexpr (K23) continue;
        // the general walk (every k; the kernels above take k = 2 and 3):
        // depth first from the block's children down to the scanned parents
        for (uint32_t ci = 0; ci < nb; ci++) {
            if (k >= 2) {
                const uint32_t c = __builtin_amdgcn_readlane(cw, (int)ci);
                step_level(task, lev(Lu - 1), lvl(Lu), c, a.s + Lu - 1, lane);
                if (lane == 0) {
                    sv.live[Lu] = live_after(lvp, c);
                    sv.idx[Lu] = j0 + ci;
                }
            }
            if (lane == 0 && Lu + 1 < MAXLEV) sv.idx[Lu + 1] = 0;
            wave_sync();
            // depth-first over positions s+Lu .. D-2, scanning position D-1
            int L = Lu;
            for (;;) {
                if (L == k - 1) {
                    JPath jpath;
                    jpath.lds = &sv.idx[1];
                    jpath.j0 = jpath.j1 = 0;
                    {
                        const TestState *PP = lvl(L);
                        const uint32_t lvL = uni(sv.live[L]);
                        const ParentSummary ps = summarize_parent(task, PP, (uint32_t)(a.D - 1), lane, my_tgt);
                        if (lane == 0) wc.c[3]++;
                        wave_sync();
                        if (ps.pm == 0u)
                            scan_parent_bits(task, a, k, cand, PP, lvL, pi, jpath, lane, wc, sv, pool, tst, ps);
                        else
                            scan_parent(task, a, k, cand, PP, lvL, pi, jpath, lane, wc, sv.cq, ps);
                    }
                    wave_sync();
                    if (L == Lu) break;
                    L--;
                    if (lane == 0) sv.idx[L + 1]++;
                    wave_sync();
                    continue;
                }
                const int pos = a.s + L;
                const int mf = a.D - pos - 1;
                const uint32_t lvL = uni(sv.live[L]);
                const int n = cand_count(cand, lvL, mf);
                const int ix = (int)uni(sv.idx[L + 1]);
                if (ix >= n) {
                    if (L == Lu) break;
                    L--;
                    if (lane == 0) sv.idx[L + 1]++;
                    wave_sync();
                    continue;
                }
                const uint32_t c = uni(cand_at(cand, lvL, mf, ix));
                step_level(task, lev(L), lvl(L + 1), c, pos, lane);
                if (lane == 0) {
                    sv.live[L + 1] = live_after(lvL, c);
                    if (L + 2 < MAXLEV) sv.idx[L + 2] = 0;
                }
                wave_sync();
                L++;
            }
        }
    }
    if (pool.qn) flush_pool();
    wave_sync();
    const unsigned long long evs = wave_sum((unsigned long long)sv.evl[lane]);
    if (lane == 0) sv.wc[2] += evs;
    if (tst) {
        const unsigned long long p2 = sv.pa2[lane];
        const unsigned long long reach = wave_sum(p2 & 0xFFFFFFFFull), ended = wave_sum(p2 >> 32);
        if (lane == 0) {
            tst[2][0] += (uint32_t)reach;
            tst[2][1] += (uint32_t)ended;
        }
    }
    wave_sync();
    if (lane < 4) atomicAdd(&lch.counters[lane], sv.wc[lane]);
    if (SHARED && lch.xcounters && lane < 2) atomicAdd(&lch.xcounters[lane], sv.wx[lane]);
    if (!MULTI && lch.tstat) {
        __syncthreads();
        for (int i = threadIdx.x; i < MAXT * 2; i += SCAN_WG) {
            uint32_t v = 0;
            for (int ww = 0; ww < SCAN_WAVES; ww++) v += (&s_wave[ww].tst[0][0])[i];
            if (v) atomicAdd(&lch.tstat[i], (unsigned long long)v);
        }
    }
}

// One-item launches: register budget of 6 waves per SIMD (at most 80
// VGPRs), measured on the is_power2_32 depth-8 window faster than 5 waves
// (96 VGPRs), 7 and 8 despite the spills (10.7 ms per launch against 11.1,
// 10.8 and 15.3).  Multi-item launches (more live context per unit): 3
// waves per SIMD (168 VGPRs, 21 spilled), measured on the server's 1,000-task
// suite at depth budget 4 against 4 waves (128 VGPRs, 95 spilled) and 2
// (188, none): 4.70 s of kernel time against 4.97 and 4.80.
__global__ __launch_bounds__(SCAN_WG, 6) void scan_kernel_one(ScanLaunch lch, ScanArgs a1, DevTask task1)
{
    scan_body<false, 0, false>(lch, a1, task1);
}

// One-item launches of 2- and 3-instruction suffixes (the block's children
// are the scanned parents): the same body without the general walk, with
// the suffix length a constant (the hit paths' per-position selects fold).
__global__ __launch_bounds__(SCAN_WG, 6) void scan_kernel_k2(ScanLaunch lch, ScanArgs a1, DevTask task1)
{
    scan_body<false, 2, false>(lch, a1, task1);
}

__global__ __launch_bounds__(SCAN_WG, 6) void scan_kernel_k3(ScanLaunch lch, ScanArgs a1, DevTask task1)
{
    scan_body<false, 3, false>(lch, a1, task1);
}

// The same with the last position shared per block (tile mode TILE_SHARED).
#ifndef PCA_SHARED_WAVES
#define PCA_SHARED_WAVES 6
#endif
__global__ __launch_bounds__(SCAN_WG, PCA_SHARED_WAVES) void scan_kernel_k2s(ScanLaunch lch, ScanArgs a1, DevTask task1)
{
    scan_body<false, 2, true>(lch, a1, task1);
}

__global__ __launch_bounds__(SCAN_WG, PCA_SHARED_WAVES) void scan_kernel_k3s(ScanLaunch lch, ScanArgs a1, DevTask task1)
{
    scan_body<false, 3, true>(lch, a1, task1);
}

__global__ __launch_bounds__(SCAN_WG, 3) void scan_kernel_multi(ScanLaunch lch, ScanArgs a1, DevTask task1)
{
    scan_body<true, 0, true>(lch, a1, task1);
}

// The kernel argument of a one-item launch: its work item with the launch's
// counters and grammar tables.
ScanArgs one_item_args(const ScanLaunch &l, const ScanItem &it)
{
    ScanArgs a;
    std::memset(&a, 0, sizeof(a));
    a.prefixes = it.prefixes;
    a.n_prefix = it.n_prefix;
    a.blocks = it.blocks;
    a.n_units = it.n_units;
    a.D = it.D;
    a.s = it.s;
    a.k = it.k;
    a.tile = it.tile;
    a.cand = l.cand;
    a.hits = it.hits;
    a.hit_cap = it.hit_cap;
    a.hit_count = it.hit_count;
    a.counters = l.counters;
    a.work = l.work;
    return a;
}

// The kernel of a one-item launch: by suffix length and tile mode.
typedef void (*ScanKernelFn)(ScanLaunch, ScanArgs, DevTask);
ScanKernelFn scan_kernel_for(const ScanItem &it)
{
    const bool shared = it.tile.mode == TILE_SHARED;
    if (it.k == 2) return shared ? scan_kernel_k2s : scan_kernel_k2;
    if (it.k == 3) return shared ? scan_kernel_k3s : scan_kernel_k3;
    return scan_kernel_one;
}

// ------------------------------------------------------------ wavefront
__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return x;
}

// Child state on one test: parent row P with candidate c at kernel position
// pos, with the exact skip / fault rules of step_test.
__device__ __forceinline__ void child_row(const TestState &p, uint32_t c, uint32_t pos, TestState &out)
{
    out = p;
    step_test(out, c, pos);
}

__device__ __forceinline__ uint64_t row_words_hash(uint64_t h, const TestState &s)
{
    const uint64_t w0 = (uint64_t)s.r[1] | ((uint64_t)s.r[2] << 16) | ((uint64_t)s.r[3] << 32) |
                        ((uint64_t)s.r[4] << 48);
    const uint64_t w1 = (uint64_t)s.r[5] | ((uint64_t)s.r[6] << 16) | ((uint64_t)s.r[7] << 32) |
                        ((uint64_t)s.flags << 48);
    h = mix64(h ^ w0);
    return mix64(h ^ w1 ^ 0x5851F42D4C957F2Dull);
}

__device__ __forceinline__ uint64_t upper_parent(const uint64_t *off, uint64_t n, uint64_t idx)
{
    uint64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (off[mid] <= idx) lo = mid; else hi = mid;
    }
    return lo;
}

__global__ void expand_count_kernel(ExpandArgs a, uint64_t *counts)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_parents) return;
    counts[i] = (uint64_t)cand_count(a.cand, a.parents[i].live, a.D - a.pos - 1);
}

// Pass 1: every non-BR child inserts its 64-bit key with atomicMin of its
// global child index (= DFS rank order), so the slot ends at the first
// visitor whatever the arrival order.
// The work loop of the insert and survive passes: units are (parent, slice)
// pairs, taken `grab` at a time; the slice covers children [jb, je) of the
// parent, whole 64-child blocks (the last one may be partial).
struct ExpandUnits {
    uint64_t next = 0, end = 0;
    uint32_t S, grab;
    uint64_t n_units;
    __device__ ExpandUnits(const ExpandArgs &a, bool sliced)
        : S(sliced && a.slices > 1u ? a.slices : 1u), grab(a.grab > 1u ? a.grab : 1u), n_units(a.n_parents * S) {}
    // false when the launch's units are exhausted
    __device__ bool take(const ExpandArgs &a, int lane, uint64_t *parent, uint32_t *slice)
    {
        if (next >= end) {
            unsigned long long g0 = 0;
            if (lane == 0) g0 = atomicAdd(a.work, (unsigned long long)grab);
            g0 = uni64(g0);
            if (g0 >= n_units) return false;
            next = g0;
            end = min((uint64_t)g0 + grab, n_units);
        }
        const uint64_t x = next++;
        *parent = x / S;
        *slice = (uint32_t)(x - *parent * S);
        return true;
    }
    __device__ void range(int n, uint32_t slice, int *jb, int *je) const
    {
        const int per = (((n + (int)S - 1) / (int)S) + 63) & ~63;
        *jb = min(n, (int)slice * per);
        *je = min(n, *jb + per);
    }
};

__global__ __launch_bounds__(SCAN_WG) void expand_insert_kernel(DevTask task, ExpandArgs a)
{
    __shared__ TestState s_st[SCAN_WAVES][MAXT];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int mf = a.D - a.pos - 1;
    ExpandUnits units(a, true);
    uint64_t u;
    uint32_t slice;
    while (units.take(a, lane, &u, &slice)) {
        const pca_prefix_t *P = &a.parents[u];
        const uint32_t live = uni(P->live);
        replay_prefix(task, P, a.pos, s_st[w], lane);
        wave_sync();
        // the parent's rows, for pass 2's confirmations (one slice writes them)
        if (slice == 0 && lane < task.T) a.parent_rows[u * MAXT + lane] = s_st[w][lane];
        const int n = cand_count(a.cand, live, mf);
        int jb, je;
        units.range(n, slice, &jb, &je);
        const uint64_t base = a.child_off[u];
        for (int j0 = jb; j0 < je; j0 += 64) {
            const int j = j0 + lane;
            if (j >= je) continue;
            const uint32_t c = cand_at(a.cand, live, mf, j);
            if (insn_op(c) == OP_BR) continue;
            uint64_t h = 0x243F6A8885A308D3ull;
            for (int t = 0; t < task.T; t++) {
                TestState cs;
                child_row(s_st[w][t], c, (uint32_t)a.pos, cs);
                h = row_words_hash(h, cs);
            }
            h = (h & a.key_mask) | 1ull;
            const unsigned long long idx = base + (unsigned long long)j;
            a.child_hash[idx] = h;   // survive reads it back instead of recomputing
            uint64_t slot = (h >> 1) & a.table_mask;
            for (uint64_t probe = 0;; probe++) {
                if (probe > a.table_mask) { atomicAdd(&a.counters[1], 1ull); break; }
                // Read-modify-write atomics are the cost of this pass, so
                // relaxed loads go first.  A key is written once (0 -> key)
                // and a slot's index only decreases, so any value a load
                // returns, however stale, is safe to act on: a nonzero key
                // is final (ours: found; another: next slot), and an index
                // at or below ours makes the atomicMin a no-op.  Only a slot
                // that reads empty needs the CAS.
                unsigned long long *k = &a.table[2 * slot];
                unsigned long long prev = __hip_atomic_load(k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (prev == 0ull) prev = atomicCAS(k, 0ull, (unsigned long long)h);
                if (prev == 0ull || prev == h) {
                    unsigned long long *mi = &a.table[2 * slot + 1];
                    if (prev == 0ull || __hip_atomic_load(mi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > idx)
                        atomicMin(mi, idx);
                    break;
                }
                slot = (slot + 1) & a.table_mask;
            }
        }
        wave_sync();
    }
}

// Pass 2: a child survives if it is a BR or the first visitor of its key
// (the key pass 1 stored per child, read back coalesced instead of stepping
// the child again on every test).
// A later visitor is confirmed against the winner's full state vector before
// it is dropped; a mismatch is a 64-bit collision and is counted (the host
// refuses the level).  Survivor bits are packed with a wave ballot.
__global__ __launch_bounds__(SCAN_WG) void expand_survive_kernel(DevTask task, ExpandArgs a)
{
    __shared__ TestState s_st[SCAN_WAVES][MAXT];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int mf = a.D - a.pos - 1;
    ExpandUnits units(a, true);
    uint64_t u;
    uint32_t slice;
    while (units.take(a, lane, &u, &slice)) {
        const pca_prefix_t *P = &a.parents[u];
        const uint32_t live = uni(P->live);
        replay_prefix(task, P, a.pos, s_st[w], lane);
        wave_sync();
        const int n = cand_count(a.cand, live, mf);
        int jb, je;
        units.range(n, slice, &jb, &je);
        const uint64_t base = a.child_off[u];
        unsigned long long kept = 0;
        for (int j0 = jb; j0 < je; j0 += 64) {
            const int j = j0 + lane;
            bool keep = false;
            if (j < je) {
                const uint32_t c = cand_at(a.cand, live, mf, j);
                if (insn_op(c) == OP_BR) {
                    keep = true;
                } else {
                    const unsigned long long idx = base + (unsigned long long)j;
                    const uint64_t h = a.child_hash[idx];   // insert's key of this child
                    uint64_t slot = (h >> 1) & a.table_mask;
                    unsigned long long win = ~0ull;
                    for (uint64_t probe = 0; probe <= a.table_mask; probe++) {
                        const unsigned long long k = a.table[2 * slot];
                        if (k == h) { win = a.table[2 * slot + 1]; break; }
                        if (k == 0ull) break;
                        slot = (slot + 1) & a.table_mask;
                    }
                    if (win == idx) {
                        keep = true;
                    } else if (win != ~0ull) {
                        // confirm: the winner's state test by test, from
                        // its parent's rows (stored by pass 1)
                        const uint64_t wp = upper_parent(a.child_off, a.n_parents, win);
                        const uint32_t wc = cand_at(a.cand, a.parents[wp].live, mf, (int)(win - a.child_off[wp]));
                        const TestState *WR = &a.parent_rows[wp * MAXT];
                        bool same = true;
                        for (int t = 0; t < task.T && same; t++) {
                            TestState ws = WR[t];
                            step_test(ws, wc, (uint32_t)a.pos);
                            TestState cs;
                            child_row(s_st[w][t], c, (uint32_t)a.pos, cs);
                            for (int r = 1; r < 8; r++) same = same && ws.r[r] == cs.r[r];
                            same = same && ws.flags == cs.flags;
                        }
                        if (!same) {
                            atomicAdd(&a.counters[0], 1ull);
                            keep = true;
                        }
                    } else {
                        atomicAdd(&a.counters[1], 1ull);
                        keep = true;
                    }
                }
            }
            const unsigned long long b = bal(keep);
            kept += (unsigned long long)__popcll(b);
            const uint64_t bit0 = base + (uint64_t)j0;
            // 64 consecutive children -> up to two 32-bit words
            if (lane == 0) {
                const uint64_t wi = bit0 >> 5;
                const uint32_t sh = (uint32_t)(bit0 & 31u);
                const uint64_t valid = (je - j0) >= 64 ? ~0ull : ((1ull << (je - j0)) - 1ull);
                const unsigned long long bv = b & valid;
                atomicOr(&a.keep_bits[wi], (uint32_t)(bv << sh));
                atomicOr(&a.keep_bits[wi + 1], (uint32_t)((bv << sh) >> 32));
                if (sh) atomicOr(&a.keep_bits[wi + 2], (uint32_t)(bv >> (64 - sh)));
            }
        }
        if (lane == 0 && kept) atomicAdd((unsigned long long *)&a.keep_count[u], kept);
        wave_sync();
    }
}

// Pass 3: order-preserving compaction.  A wave owns a parent; the output
// slot of a survivor is the parent's base plus a ballot prefix count.
__global__ __launch_bounds__(SCAN_WG) void expand_compact_kernel(ExpandArgs a)
{
    const int lane = threadIdx.x & 63;
    const int mf = a.D - a.pos - 1;
    ExpandUnits units(a, false);
    uint64_t u;
    uint32_t slice;
    while (units.take(a, lane, &u, &slice)) {
        const pca_prefix_t P = a.parents[u];
        const int n = cand_count(a.cand, P.live, mf);
        const uint64_t base = a.child_off[u];
        uint64_t outp = a.keep_off[u];
        for (int j0 = 0; j0 < n; j0 += 64) {
            const int j = j0 + lane;
            bool keep = false;
            if (j < n) {
                const uint64_t bit = base + (uint64_t)j;
                keep = (a.keep_bits[bit >> 5] >> (bit & 31u)) & 1u;
            }
            const unsigned long long b = bal(keep);
            if (keep) {
                const unsigned long long below = b & ((1ull << lane) - 1ull);
                pca_prefix_t c = P;
                const uint32_t insn = cand_at(a.cand, P.live, mf, j);
                c.insns[a.pos] = (uint16_t)insn;
                c.num_insns = (uint8_t)(a.pos + 1);
                c.live = (uint8_t)live_after(P.live, insn);
                a.out[outp + (uint64_t)__popcll(below)] = c;
            }
            outp += (uint64_t)__popcll(b);
        }
    }
}

// ---------------------------------------------------------- prefix sums
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_BLOCK = 256;
constexpr uint64_t SCAN_TILE = (uint64_t)SCAN_ITEMS * SCAN_BLOCK;

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t v, int lane)
{
    for (int o = 1; o < 64; o <<= 1) {
        const uint64_t t = __shfl_up(v, o);
        if (lane >= o) v += t;
    }
    return v;
}

// block-wide exclusive scan of one value per thread; returns the block total
__device__ uint64_t block_excl_scan(uint64_t v, uint64_t *out_excl)
{
    __shared__ uint64_t s_w[SCAN_BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t inc = wave_incl_scan(v, lane);
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint64_t wbase = 0, total = 0;
    for (int i = 0; i < SCAN_BLOCK / 64; i++) {
        if (i < w) wbase += s_w[i];
        total += s_w[i];
    }
    *out_excl = wbase + inc - v;
    __syncthreads();
    return total;
}

__global__ void scan_tile_sums(const uint64_t *data, uint64_t n, uint64_t *sums)
{
    const uint64_t t0 = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; i++)
        if (t0 + i < n) s += data[t0 + i];
    uint64_t ex;
    const uint64_t tot = block_excl_scan(s, &ex);
    if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

__global__ void scan_tile_apply(uint64_t *data, uint64_t n, const uint64_t *sums_excl)
{
    const uint64_t t0 = (uint64_t)blockIdx.x * SCAN_TILE + (uint64_t)threadIdx.x * SCAN_ITEMS;
    uint64_t vals[SCAN_ITEMS];
    uint64_t s = 0;
    for (int i = 0; i < SCAN_ITEMS; i++) {
        vals[i] = (t0 + i < n) ? data[t0 + i] : 0;
        s += vals[i];
    }
    uint64_t ex;
    block_excl_scan(s, &ex);
    uint64_t run = ex + (sums_excl ? sums_excl[blockIdx.x] : 0);
    for (int i = 0; i < SCAN_ITEMS; i++) {
        if (t0 + i < n) data[t0 + i] = run;
        run += vals[i];
    }
}

// ------------------------------------------------------------ state match
__device__ __forceinline__ bool set_has(const unsigned long long *t, uint64_t mask, uint64_t key)
{
    uint64_t slot = (key >> 1) & mask;
    for (uint64_t probe = 0; probe <= mask; probe++) {
        const unsigned long long k = t[slot];
        if (k == key) return true;
        if (k == 0ull) return false;
        slot = (slot + 1) & mask;
    }
    return false;
}

__device__ __forceinline__ uint32_t map_get(const unsigned long long *t, uint64_t mask, uint64_t key)
{
    uint64_t slot = (key >> 1) & mask;
    for (uint64_t probe = 0; probe <= mask; probe++) {
        const unsigned long long k = t[2 * slot];
        if (k == key) return (uint32_t)t[2 * slot + 1];
        if (k == 0ull) return 0xFFFFFFFFu;
        slot = (slot + 1) & mask;
    }
    return 0xFFFFFFFFu;
}

// Every node at kernel position p = s + kk - 1 (all prefixes, all
// suffixes, BR children excluded: they never enter an OEP table) whose
// state hashes like a query state.  Test 0's row filters first; only its
// survivors hash all tests.
__global__ __launch_bounds__(SCAN_WG) void match_kernel(DevTask task, MatchArgs a)
{
    __shared__ TestState s_st[SCAN_WAVES][MAXLEV][MAXT];
    __shared__ uint32_t s_live[SCAN_WAVES][MAXLEV];
    __shared__ uint32_t s_idx[SCAN_WAVES][MAXLEV];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int k = a.kk;
    const int Lu = k <= 1 ? 0 : (k == 2 ? 1 : 2);
    for (;;) {
        unsigned long long u = 0;
        if (lane == 0) u = atomicAdd(a.work, 1ull);
        u = __shfl(u, 0);
        if (u >= a.n_units) break;
        uint32_t pi, js = 0, js1 = 0;
        if (k >= 2) {
            uint32_t lo = 0, hi = a.n_groups;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (a.group_off[mid] <= u) lo = mid; else hi = mid;
            }
            const ScanGroup G = a.groups[lo];
            pi = uni(G.prefix);
            const uint32_t j = uni((uint32_t)(u - a.group_off[lo]));
            if (k == 2) js = j;
            else { js = uni(G.js); js1 = j; }
        } else {
            pi = uni((uint32_t)u);
        }
        const pca_prefix_t *P = &a.prefixes[pi];
        const uint32_t live0 = uni(P->live);
        replay_prefix(task, P, a.s, s_st[w][0], lane);
        if (lane == 0) s_live[w][0] = live0;
        wave_sync();
        if (k >= 2) {
            const uint32_t c1 = uni(cand_at(a.cand, live0, a.D - a.s - 1, (int)js));
            step_level(task, s_st[w][0], s_st[w][1], c1, a.s, lane);
            if (lane == 0) { s_live[w][1] = live_after(live0, c1); s_idx[w][1] = js; }
            wave_sync();
            if (k >= 3) {
                const uint32_t lv1 = uni(s_live[w][1]);
                const uint32_t c2 = uni(cand_at(a.cand, lv1, a.D - a.s - 2, (int)js1));
                step_level(task, s_st[w][1], s_st[w][2], c2, a.s + 1, lane);
                if (lane == 0) { s_live[w][2] = live_after(lv1, c2); s_idx[w][2] = js1; }
                wave_sync();
            }
        }
        if (lane == 0 && Lu + 1 < MAXLEV) s_idx[w][Lu + 1] = 0;
        wave_sync();
        int L = Lu;
        for (;;) {
            if (L == k - 1) {
                // lanes = candidates at the last position p
                const int pos = a.s + L;
                const int mf = a.D - pos - 1;
                const uint32_t lv = uni(s_live[w][L]);
                const int n = cand_count(a.cand, lv, mf);
                const TestState *PP = s_st[w][L];
                for (int base = 0; base < n; base += 64) {
                    const int j = base + lane;
                    if (j < n) {
                        const uint32_t c = cand_at(a.cand, lv, mf, j);
                        if (insn_op(c) != OP_BR) {
                            TestState c0 = PP[0];
                            step_test(c0, c, (uint32_t)pos);
                            const uint64_t h0 = row_words_hash(0x243F6A8885A308D3ull, c0) | 1ull;
                            if (set_has(a.q0, a.q0_mask, h0)) {
                                uint64_t h = 0x243F6A8885A308D3ull;
                                for (int t = 0; t < task.T; t++) {
                                    TestState ct = PP[t];
                                    step_test(ct, c, (uint32_t)pos);
                                    h = row_words_hash(h, ct);
                                }
                                const uint32_t q = map_get(a.qf, a.qf_mask, h | 1ull);
                                if (q != 0xFFFFFFFFu) {
                                    const uint32_t slot = atomicAdd(a.count, 1u);
                                    if (slot < a.cap) {
                                        MatchRec r;
                                        r.prefix = pi;
                                        r.query = q;
                                        for (int qq = 0; qq < PCAGPU_MAX_KDEPTH; qq++) r.j[qq] = 0;
                                        for (int qq = 1; qq <= L; qq++) r.j[qq - 1] = (uint16_t)s_idx[w][qq];
                                        r.j[L] = (uint16_t)j;
                                        a.out[slot] = r;
                                    }
                                }
                            }
                        }
                    }
                }
                wave_sync();
                if (L == Lu) break;
                L--;
                if (lane == 0) s_idx[w][L + 1]++;
                wave_sync();
                continue;
            }
            const int pos = a.s + L;
            const int mf = a.D - pos - 1;
            const uint32_t lvL = uni(s_live[w][L]);
            const int n = cand_count(a.cand, lvL, mf);
            const int ix = (int)uni(s_idx[w][L + 1]);
            if (ix >= n) {
                if (L == Lu) break;
                L--;
                if (lane == 0) s_idx[w][L + 1]++;
                wave_sync();
                continue;
            }
            const uint32_t c = uni(cand_at(a.cand, lvL, mf, ix));
            step_level(task, s_st[w][L], s_st[w][L + 1], c, pos, lane);
            if (lane == 0) {
                s_live[w][L + 1] = live_after(lvL, c);
                if (L + 2 < MAXLEV) s_idx[w][L + 2] = 0;
            }
            wave_sync();
            L++;
        }
    }
}

// Exhaustive ALU sweep: sum over a in [0, 2^16), b in [b_lo, b_hi) of a
// 64-bit mix of (a, b, value, flags, fault) with the device ALU semantics.
__device__ __forceinline__ uint64_t dg_mix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__global__ void alu_digest_kernel(uint32_t op, uint32_t b_lo, uint32_t b_hi, unsigned long long *out)
{
    const uint32_t b = b_lo + blockIdx.x;
    if (b >= b_hi) return;
    unsigned long long h = 0;
    for (uint32_t a = threadIdx.x; a < 65536u; a += blockDim.x) {
        uint32_t f = 0, v;
        bool fault = false;
        if (op == OP_CMP) v = 0, f = flags_sub(a, b);
        else v = alu_eval(op, a, b, 0, 0, &f, &fault);
        if (fault) v = 0, f = 0;
        h += dg_mix((uint64_t)a | ((uint64_t)b << 16) | ((uint64_t)v << 32) | ((uint64_t)f << 48) |
                    ((uint64_t)fault << 56));
    }
    h = wave_sum(h);
    if ((threadIdx.x & 63) == 0) atomicAdd(out, h);
}

__global__ void zero_words_kernel(uint32_t *p0, uint32_t n0, uint32_t *p1, uint32_t n1, uint32_t *p2, uint32_t n2,
                                  uint32_t *p3, uint32_t n3)
{
    const uint32_t i = threadIdx.x;
    if (i < n0) p0[i] = 0u;
    if (i < n1) p1[i] = 0u;
    if (i < n2) p2[i] = 0u;
    if (i < n3) p3[i] = 0u;
}

__global__ void table_init_kernel(unsigned long long *t, uint64_t cap)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < cap) {
        t[2 * i] = 0ull;
        t[2 * i + 1] = ~0ull;
    }
}

}  // namespace

// ------------------------------------------------------------- launchers
void launch_match(const DevTask &task, const MatchArgs &a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(match_kernel, dim3(grid), dim3(SCAN_WG), 0, st, task, a);
}

void launch_alu_digest(uint32_t op, uint32_t b_lo, uint32_t b_hi, unsigned long long *out, hipStream_t st)
{
    if (b_hi > b_lo)
        hipLaunchKernelGGL(alu_digest_kernel, dim3(b_hi - b_lo), dim3(256), 0, st, op, b_lo, b_hi, out);
}

void launch_table_init(unsigned long long *table, uint64_t cap, hipStream_t st)
{
    hipLaunchKernelGGL(table_init_kernel, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, st, table, cap);
}

size_t scan_dyn_lds(uint32_t levels)
{
    const size_t per_wave = (levels > 2u ? levels : 2u) - 1u;
    return (size_t)SCAN_WAVES * per_wave * MAXT * sizeof(TestState);
}

hipError_t scan_occupancy(bool multi, uint32_t levels, int *blocks_per_cu)
{
    const size_t lds = scan_dyn_lds(levels);
    if (multi) return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, scan_kernel_multi, SCAN_WG, lds);
    // a one-item launch's levels are its suffix length k (k >= 1)
    if (levels == 2u || levels == 3u) {
        // the shared and per-child kernels of this length: the smaller
        // occupancy of the two sizes the persistent grid
        int b0 = 0, b1 = 0;
        hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &b0, levels == 2u ? scan_kernel_k2s : scan_kernel_k3s, SCAN_WG, lds);
        if (e != hipSuccess) return e;
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, levels == 2u ? scan_kernel_k2 : scan_kernel_k3,
                                                         SCAN_WG, lds);
        *blocks_per_cu = b0 < b1 ? b0 : b1;
        return e;
    }
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, scan_kernel_one, SCAN_WG, lds);
}

hipError_t aux_occupancy(AuxKernel which, int *blocks_per_cu)
{
    switch (which) {
    case AuxKernel::ExpandInsert:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, expand_insert_kernel, SCAN_WG, 0);
    case AuxKernel::ExpandSurvive:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, expand_survive_kernel, SCAN_WG, 0);
    case AuxKernel::ExpandCompact:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, expand_compact_kernel, SCAN_WG, 0);
    case AuxKernel::Match:
        return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, match_kernel, SCAN_WG, 0);
    default:
        return hipErrorInvalidValue;
    }
}

void scan_node_params(const ScanLaunch &l, const ScanItem &it, const DevTask &task, int grid, ScanNodeArgs *args,
                      hipKernelNodeParams *p)
{
    args->l = l;
    args->a1 = one_item_args(l, it);
    args->task = task;
    args->ptrs[0] = &args->l;
    args->ptrs[1] = &args->a1;
    args->ptrs[2] = &args->task;
    p->func = reinterpret_cast<void *>(scan_kernel_for(it));
    p->gridDim = dim3((unsigned)grid);
    p->blockDim = dim3(SCAN_WG);
    p->sharedMemBytes = (unsigned)scan_dyn_lds(l.levels);
    p->kernelParams = args->ptrs;
    p->extra = nullptr;
}

void zero_node_params(uint32_t *const ptr[4], const uint32_t n[4], ZeroNodeArgs *args, hipKernelNodeParams *p)
{
    args->p0 = ptr[0];
    args->p1 = ptr[1];
    args->p2 = ptr[2];
    args->p3 = ptr[3];
    args->n0 = n[0];
    args->n1 = n[1];
    args->n2 = n[2];
    args->n3 = n[3];
    args->ptrs[0] = &args->p0;
    args->ptrs[1] = &args->n0;
    args->ptrs[2] = &args->p1;
    args->ptrs[3] = &args->n1;
    args->ptrs[4] = &args->p2;
    args->ptrs[5] = &args->n2;
    args->ptrs[6] = &args->p3;
    args->ptrs[7] = &args->n3;
    p->func = reinterpret_cast<void *>(zero_words_kernel);
    p->gridDim = dim3(1);
    p->blockDim = dim3(256);
    p->sharedMemBytes = 0;
    p->kernelParams = args->ptrs;
    p->extra = nullptr;
}

void launch_scan(const ScanLaunch &l, const ScanItem *host_items, const DevTask &task0, int grid,
                 hipStream_t st)
{
    const size_t lds = scan_dyn_lds(l.levels);
    const ScanArgs a1 = one_item_args(l, host_items[0]);
    if (l.n_items == 1)
        hipLaunchKernelGGL(scan_kernel_for(host_items[0]), dim3(grid), dim3(SCAN_WG), lds, st, l, a1, task0);
    else
        hipLaunchKernelGGL(scan_kernel_multi, dim3(grid), dim3(SCAN_WG), lds, st, l, a1, task0);
}

void launch_expand_count(const ExpandArgs &a, uint64_t *counts, hipStream_t st)
{
    const uint64_t blocks = (a.n_parents + 255) / 256;
    if (blocks) hipLaunchKernelGGL(expand_count_kernel, dim3((unsigned)blocks), dim3(256), 0, st, a, counts);
}

void launch_expand_insert(const DevTask &task, const ExpandArgs &a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(expand_insert_kernel, dim3(grid), dim3(SCAN_WG), 0, st, task, a);
}

void launch_expand_survive(const DevTask &task, const ExpandArgs &a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(expand_survive_kernel, dim3(grid), dim3(SCAN_WG), 0, st, task, a);
}

void launch_expand_compact(const ExpandArgs &a, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(expand_compact_kernel, dim3(grid), dim3(SCAN_WG), 0, st, a);
}

uint64_t exclusive_scan_tmp_words(uint64_t n)
{
    // one level of tile sums per recursion step, plus slack
    uint64_t words = 0, m = n + 1;
    while (m > SCAN_TILE) {
        m = (m + SCAN_TILE - 1) / SCAN_TILE;
        words += m + 1;
    }
    return words + 1;
}

// Exclusive scan of data[0..n] (n+1 values, the last one is usually 0 and
// receives the grand total).
void launch_exclusive_scan(uint64_t *data, uint64_t n, uint64_t *tmp, hipStream_t st)
{
    const uint64_t len = n + 1;
    const uint64_t tiles = (len + SCAN_TILE - 1) / SCAN_TILE;
    if (tiles <= 1) {
        hipLaunchKernelGGL(scan_tile_apply, dim3(1), dim3(SCAN_BLOCK), 0, st, data, len, (const uint64_t *)nullptr);
        return;
    }
    uint64_t *sums = tmp;
    hipLaunchKernelGGL(scan_tile_sums, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, st, data, len, sums);
    launch_exclusive_scan(sums, tiles - 1, tmp + tiles + 1, st);
    hipLaunchKernelGGL(scan_tile_apply, dim3((unsigned)tiles), dim3(SCAN_BLOCK), 0, st, data, len, sums);
}

}  // namespace pca
