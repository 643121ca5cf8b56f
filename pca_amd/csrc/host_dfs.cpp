// host_dfs.cpp -- exact-OEP prefix producer (see host_dfs.h).
#include "host_dfs.h"

#include <cstring>

namespace pca {

namespace {
inline uint64_t key_hash(const uint8_t *p, int n)
{
    // 64-bit multiply-xorshift over 8-byte words; exactness comes from the
    // byte compare, the hash only spreads slots.
    uint64_t h = 0x9E3779B97F4A7C15ull ^ (uint64_t)n;
    int i = 0;
    for (; i + 8 <= n; i += 8) {
        uint64_t w;
        std::memcpy(&w, p + i, 8);
        h = (h ^ w) * 0xBF58476D1CE4E5B9ull;
        h ^= h >> 31;
    }
    uint64_t tail = 0;
    std::memcpy(&tail, p + i, (size_t)(n - i));
    h = (h ^ tail) * 0x94D049BB133111EBull;
    h ^= h >> 29;
    return h | 1;   // 0 marks an empty slot
}
}  // namespace

void ExactKeySet::reset(int klen)
{
    klen_ = klen;
    hash_.clear();
    slot_.clear();
    arena_.clear();
    n_ = 0;
}

void ExactKeySet::clear()
{
    std::fill(hash_.begin(), hash_.end(), 0);
    arena_.clear();
    n_ = 0;
}

void ExactKeySet::grow()
{
    size_t cap = hash_.empty() ? 4096 : hash_.size() * 2;
    std::vector<uint64_t> nh(cap, 0);
    std::vector<uint32_t> ns(cap, 0);
    for (size_t i = 0; i < hash_.size(); i++) {
        if (!hash_[i]) continue;
        size_t j = (hash_[i] >> 1) & (cap - 1);
        while (nh[j]) j = (j + 1) & (cap - 1);
        nh[j] = hash_[i];
        ns[j] = slot_[i];
    }
    hash_.swap(nh);
    slot_.swap(ns);
}

int ExactKeySet::insert(const uint8_t *key)
{
    if ((n_ + 1) * 2 > hash_.size()) grow();
    const uint64_t h = key_hash(key, klen_);
    const size_t mask = hash_.size() - 1;
    size_t j = (h >> 1) & mask;
    while (hash_[j]) {
        if (hash_[j] == h && std::memcmp(&arena_[(size_t)slot_[j] * klen_], key, (size_t)klen_) == 0)
            return 0;
        j = (j + 1) & mask;
    }
    hash_[j] = h;
    slot_[j] = (uint32_t)n_;
    arena_.insert(arena_.end(), key, key + klen_);
    n_++;
    return 1;
}

namespace {
struct Walker {
    const pca_task_t *task;
    int D, s, T;
    const std::function<bool(const pca_prefix_t &)> *emit;
    DfsStats *st;
    std::vector<ExactKeySet> sets;
    std::vector<TestState> states;   // [(level+1) * T], level 0 = root
    uint16_t insns[PCAGPU_MAX_KDEPTH];
    bool stop = false;

    void walk(int pos, uint32_t live)
    {
        if (stop) return;
        if (pos == s) {
            pca_prefix_t p;
            std::memset(&p, 0, sizeof(p));
            for (int i = 0; i < s; i++) p.insns[i] = insns[i];
            p.num_insns = (uint8_t)s;
            p.live = (uint8_t)live;
            st->emitted++;
            if (!(*emit)(p)) stop = true;
            return;
        }
        uint16_t cands[CAND_LMAX + 64];
        int nc = gen_candidates(cands, CAND_LMAX + 64, live, pos, D);
        const TestState *parent = &states[(size_t)pos * T];
        TestState *child = &states[(size_t)(pos + 1) * T];
        uint8_t key[PCAGPU_MAX_TESTS * 17];
        for (int i = 0; i < nc && !stop; i++) {
            const uint32_t c = cands[i];
            for (int k = 0; k < T; k++) {
                child[k] = parent[k];
                step_test(child[k], c, (uint32_t)pos);
            }
            // tools/enumerate.c:435: BR children are exempt from OEP
            if (insn_op(c) != OP_BR) {
                oep_key_from_states(child, T, key);
                if (!sets[pos].insert(key)) { st->pruned++; continue; }
            }
            st->nodes++;
            insns[pos] = (uint16_t)c;
            walk(pos + 1, live_after(live, c));
        }
    }
};
}  // namespace

void dfs_prefixes(const pca_task_t *task, int D, int s,
                  const std::function<bool(const pca_prefix_t &)> &emit, DfsStats *stats)
{
    Walker w;
    w.task = task;
    w.D = D;
    w.s = s;
    w.T = task->num_tests;
    w.emit = &emit;
    w.st = stats;
    w.sets.resize((size_t)(s > 0 ? s : 1));
    for (auto &set : w.sets) set.reset(w.T * 17 > 0 ? w.T * 17 : 1);
    w.states.resize((size_t)(s + 1) * (w.T > 0 ? w.T : 1));
    root_states(task, w.states.data());
    uint32_t live = 0;
    for (int i = 0; i < task->num_inputs && i < 7; i++) live |= 1u << (i + 1);
    w.walk(0, live);
}

}  // namespace pca
