// complete.cpp -- the reference's branchless search state and its exhaustive
// leaf completion (src/search_state.c:15-453) behind include/pcagpu.h.
//
// The reference completes a state by IDDFS: for target = 1, 2, .. it walks
// the branchless grammar depth first, checks every child for a live register
// equal to the target on every test, and prunes a child that repeats the
// register/flag state of an earlier child at the same depth (per-depth
// tables cleared per target).  The first solving child found is the answer.
//
// On the GPU the same answer comes from levels instead of a walk.  Level r
// scans every child of the level-(r-1) frontier F(r-1) for a solving one
// (the last-position scan kernel, candidates from the branchless tables);
// if there is none, F(r) = the children of F(r-1) with the first visitor of
// each state kept (the wavefront expand kernels, exact keys, rank order).
// The DFS at target r visits exactly F(1) .. F(r-1) (pruning at a depth only
// depends on the shallower ones, and the first visitor in DFS order is the
// lowest rank), so the first solving child in DFS order is the lowest
// (frontier index, candidate index) hit of the first level that has any.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstring>
#include <vector>

#include "pcagpu_impl.h"

using namespace pca;

static_assert(sizeof(pca_search_state_t) == 546, "search_state_t layout (src/search_state.h:19-25)");

namespace {

// src/search_state.c:49-187 for one instruction on every test, in the
// handle's VM arithmetic (alu_eval: the flag rules of src/vm.c).
bool state_step(const pca_task_t *task, const pca_search_state_t *src, uint16_t insn, pca_search_state_t *dst)
{
    const uint32_t op = insn_op(insn), rd = insn_rd(insn), rs = insn_rs(insn), rt = insn_rt(insn);
    switch (op) {
    case OP_ADD: case OP_SUB: case OP_MUL: case OP_DIV: case OP_MOD: case OP_AND:
    case OP_OR: case OP_XOR: case OP_SHL: case OP_SHR: case OP_ASR: case OP_LDI: case OP_ADDI:
        break;
    default:
        return false;
    }
    pca_search_state_t out = *src;
    out.depth = (uint8_t)(src->depth + 1);
    for (int t = 0; t < task->num_tests; t++) {
        uint16_t *r = out.regs[t];
        uint32_t f = out.flags[t];
        bool fault = false;
        const uint32_t a = op == OP_ADDI ? r[rd] : r[rs];
        const uint32_t v = alu_eval(op, a, r[rt], insn_imm(insn), f, &f, &fault);
        if (fault) return false;
        r[rd] = (uint16_t)v;
        out.flags[t] = (uint8_t)f;
        r[0] = 0;
    }
    if (rd != 0) out.live = (uint8_t)(src->live | (1u << rd));
    *dst = out;
    return true;
}

uint8_t leaf_outputs(const pca_task_t *task, const pca_search_state_t *s)
{
    uint8_t m = 0;
    for (int r = 1; r < 8; r++) {
        if (!(s->live >> r & 1u)) continue;
        bool ok = true;
        for (int t = 0; t < task->num_tests && ok; t++) ok = s->regs[t][r] == task->tests[t].out[0];
        if (ok) m |= (uint8_t)(1u << r);
    }
    return m;
}

// The scan of level r: every child of the n frontier records in
// d_front[cur] (r-1 instructions each below the start state), one work unit
// per record, hits into d_hits (repeated with a larger buffer on overflow).
int scan_level(pcagpu *h, const DevTask &dt, const CandDev &cd, int cur, uint64_t n, int r, uint32_t *hits_out,
               uint64_t *children_out)
{
    ScanItem it;
    ScanLaunch l;
    unsigned long long cnt[4];
    uint32_t hits = 0;
    for (;;) {
        std::memset(&it, 0, sizeof(it));
        std::memset(&l, 0, sizeof(l));
        it.task = h->d_task.p;   // read by multi-item launches only
        it.prefixes = h->d_front[cur].p;
        it.n_prefix = (uint32_t)n;
        it.n_units = n;
        it.D = r;
        it.s = r - 1;
        it.k = 1;
        it.tile = ScanTile{64, 2, TILE_SHARED};
        it.hits = h->d_hits.p;
        it.hit_count = h->d_hit_count.p;
        it.hit_cap = h->hit_cap;
        l.items = h->d_items.p;
        l.n_items = 1;
        l.levels = scan_levels(1);
        l.n_units = n;
        l.cand = cd;
        l.counters = h->d_counters.p;
        l.work = h->d_work.p;
        l.tstat = nullptr;
        const int grid = grid_for(h, n, false, l.levels);
        // a frontier record is a cheap unit (one parent): take runs of them,
        // about 8 runs per wave
        const uint64_t waves = (uint64_t)grid * SCAN_WAVES;
        l.grab = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(n / (waves * 8), 1), 32);
        HIP_TRY(hipMemsetAsync(h->d_work.p, 0, sizeof(unsigned long long), h->stream));
        HIP_TRY(hipMemsetAsync(h->d_counters.p, 0, 4 * sizeof(unsigned long long), h->stream));
        HIP_TRY(hipMemsetAsync(h->d_hit_count.p, 0, sizeof(uint32_t), h->stream));
        HIP_TRY(hipEventRecord(h->ev0, h->stream));
        launch_scan(l, &it, dt, grid, h->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(h->ev1, h->stream));
        HIP_TRY_S(h->stream, hipMemcpyAsync(cnt, h->d_counters.p, sizeof(cnt), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY_S(h->stream, hipMemcpyAsync(&hits, h->d_hit_count.p, sizeof(hits), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY_S(h->stream, hipStreamSynchronize(h->stream));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->stats.kernel_ns += (uint64_t)((double)ms * 1e6);
        h->stats.launches++;
        if (hits <= h->hit_cap) break;
        const uint32_t want = hits + hits / 2 + 1024;
        h->d_hits.release();
        h->hit_cap = 0;   // until the larger buffer exists: launches write no records
        int rc;
        if ((rc = h->d_hits.ensure(want))) return rc;
        h->hit_cap = want;
    }
    h->stats.suffixes += cnt[1];
    h->stats.tests_evaluated += cnt[2];
    h->stats.parents += cnt[3];
    h->stats.raw_hits += hits;
    *hits_out = hits;
    *children_out = cnt[1];
    return 0;
}

}  // namespace

extern "C" {

void pcagpu_state_init_root(const pca_task_t *task, pca_search_state_t *s)
{
    if (!task || !s) return;
    std::memset(s, 0, sizeof(*s));
    const int T = std::min(std::max(task->num_tests, 0), PCAGPU_MAX_TESTS);
    const int n_in = std::min(std::max(task->num_inputs, 0), 7);
    for (int t = 0; t < T; t++)
        for (int i = 0; i < n_in; i++) s->regs[t][i + 1] = task->tests[t].in[i];
    for (int i = 0; i < n_in; i++) s->live |= (uint8_t)(1u << (i + 1));
}

int pcagpu_state_step(const pca_task_t *task, const pca_search_state_t *src, uint16_t insn, pca_search_state_t *dst)
{
    if (!task || !src || !dst || task->num_tests < 0 || task->num_tests > PCAGPU_MAX_TESTS) return 0;
    return state_step(task, src, insn, dst) ? 1 : 0;
}

uint8_t pcagpu_state_leaf_outputs(const pca_task_t *task, const pca_search_state_t *s)
{
    if (!task || !s || task->num_tests < 0 || task->num_tests > PCAGPU_MAX_TESTS) return 0;
    return leaf_outputs(task, s);
}

int pcagpu_gen_candidates_branchless(uint16_t *out, int max, uint8_t live)
{
    if (!out || max < 0) return PCAGPU_EINVAL;
    return gen_candidates_branchless(out, max, live);
}

uint64_t pcagpu_complete_states_explored(pcagpu_t *h) { return h ? h->complete_explored : 0; }

int pcagpu_complete(pcagpu_t *h, const pca_search_state_t *start, int max_depth, uint16_t *solution_insns,
                    int *solution_len)
{
    if (!h || !start || !solution_insns || !solution_len) return PCAGPU_EINVAL;
    *solution_len = 0;
    (void)hipSetDevice(h->device);
    const int rc0 = settle_scan(h);   // a scan in flight shares the stream and buffers
    if (rc0) return rc0;
    h->complete_explored = 0;
    if (leaf_outputs(&h->task, start)) return 1;   // the start state already solves
    int max_r = max_depth - (int)start->depth;
    if (max_r <= 0) return 0;
    max_r = std::min(max_r, 16);
    if (max_r > PCAGPU_MAX_KDEPTH) return PCAGPU_EINVAL;   // frontier records hold 8 instructions
    HIP_TRY(hipSetDevice(h->device));
    int rc;
    if (!h->d_grammar_bl && !(h->d_grammar_bl = shared_grammar_dev(h->device, true))) return PCAGPU_ENOMEM;
    const CandDev cd = cand_dev(h->d_grammar_bl);
    const CandTable &ct = shared_grammar_branchless();
    // the start state as the root of every test, in the device test order;
    // no cycle budget (the search state has none)
    DevTask dt = h->dt;
    dt.max_cycles = INT_MAX;
    for (int t = 0; t < h->T; t++) {
        TestState &s = dt.root[t];
        std::memset(&s, 0, sizeof(s));
        std::memcpy(s.r, start->regs[h->order[t]], sizeof(s.r));
        s.flags = start->flags[h->order[t]];
    }
    pca_prefix_t root;
    std::memset(&root, 0, sizeof(root));
    root.live = start->live;
    if ((rc = h->d_front[0].ensure(1))) return rc;
    HIP_TRY(hipMemcpy(h->d_front[0].p, &root, sizeof(root), hipMemcpyHostToDevice));
    uint64_t n = 1;
    int cur = 0;
    for (int r = 1; r <= max_r && n > 0; r++) {
        uint32_t hits = 0;
        uint64_t children = 0;
        if ((rc = scan_level(h, dt, cd, cur, n, r, &hits, &children))) return rc;
        h->complete_explored += children;
        if (hits) {
            std::vector<HitRec> buf(hits);
            HIP_TRY(hipMemcpy(buf.data(), h->d_hits.p, (size_t)hits * sizeof(HitRec), hipMemcpyDeviceToHost));
            std::sort(buf.begin(), buf.end(), [](const HitRec &x, const HitRec &y) {
                return x.prefix != y.prefix ? x.prefix < y.prefix : x.j[0] < y.j[0];
            });
            // a hit is a child with a target-matching register; it solves if
            // that register is live (a register the start state left unwritten
            // can match without counting)
            uint32_t cached = UINT32_MAX;
            pca_prefix_t P;
            for (const HitRec &x : buf) {
                if (x.prefix != cached) {
                    HIP_TRY(hipMemcpy(&P, h->d_front[cur].p + x.prefix, sizeof(P), hipMemcpyDeviceToHost));
                    cached = x.prefix;
                }
                const uint16_t insn = ct.at(P.live, 0, x.j[0]);
                if (!(x.mask & live_after(P.live, insn) & 0xFEu)) continue;
                uint16_t path[PCAGPU_MAX_KDEPTH];
                for (int q = 0; q < r - 1; q++) path[q] = P.insns[q];
                path[r - 1] = insn;
                // re-verify on the host before returning it
                pca_search_state_t s = *start, nx;
                for (int q = 0; q < r; q++) {
                    if (!state_step(&h->task, &s, path[q], &nx)) return PCAGPU_EVERIFY;
                    s = nx;
                }
                if (!leaf_outputs(&h->task, &s)) return PCAGPU_EVERIFY;
                std::memcpy(solution_insns, path, (size_t)r * sizeof(uint16_t));
                *solution_len = r;
                return 1;
            }
        }
        if (r == max_r) break;
        // the same children again, now kept or dropped by state (the scan
        // above already counted them as explored)
        uint64_t kept = 0;
        if ((rc = expand_level(h, dt, cd, cur, n, r, r - 1, &kept))) return rc;
        if (kept > UINT32_MAX) return PCAGPU_EOVERFLOW;   // hit records index the frontier with 32 bits
        n = kept;
        cur ^= 1;
    }
    return 0;
}

}  // extern "C"
