// pcagpu.cpp -- host driver behind include/pcagpu.h.
//
// Replaces the host<->device seam of gpu/hybrid.cu (constant upload
// :592-621, batch emit/flush :470-521, hit download and CPU re-verification
// :784-816) and the level loop of gpu/wavefront.cu (:801-864), with one HIP
// stream per handle, pinned staging buffers and exact OEP resolution of the
// device hits on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "host_core.h"
#include "host_dfs.h"
#include "kernels.h"

using namespace pca;

namespace {

#define HIP_TRY(expr)                                  \
    do {                                               \
        if ((expr) != hipSuccess) return PCAGPU_EHIP;  \
    } while (0)

struct HostHit {
    uint64_t prefix;     // index into the depth's prefix list
    uint8_t mask, flags;
    uint16_t j[PCAGPU_MAX_KDEPTH];
};

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    int ensure(size_t want)
    {
        if (want <= n) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        if (hipMalloc((void **)&p, want * sizeof(T)) != hipSuccess) { p = nullptr; return PCAGPU_ENOMEM; }
        n = want;
        return 0;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

struct pcagpu {
    pca_task_t task;
    int T = 0;
    DevTask dt;
    CandTable ct;
    int device = 0;
    int cu_count = 256;
    std::string devname;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    pcagpu_opts_t opts;
    uint32_t waves_per_cu = 20;

    DevBuf<uint16_t> d_lists;
    DevBuf<int> d_count, d_split;
    DevBuf<pca_prefix_t> d_batch;
    DevBuf<ScanGroup> d_groups;
    DevBuf<uint64_t> d_goff;
    DevBuf<HitRec> d_hits;
    DevBuf<uint32_t> d_hit_count;
    DevBuf<unsigned long long> d_counters, d_work;
    pca_prefix_t *h_batch = nullptr;
    ScanGroup *h_groups = nullptr;
    uint64_t *h_goff = nullptr;
    uint32_t batch_cap = 0;
    uint32_t hit_cap = 0;

    // wavefront buffers
    DevBuf<pca_prefix_t> d_front[2];
    DevBuf<uint64_t> d_child_off, d_keep_cnt, d_scan_tmp;
    DevBuf<unsigned long long> d_table;
    DevBuf<uint32_t> d_keep_bits;

    // depth state
    bool in_depth = false;
    int D = 0, s = 0, k = 0;
    uint32_t batch_n = 0;
    uint64_t batch_first = 0;   // depth-prefix index of h_batch[0]
    std::vector<pca_prefix_t> prefixes;
    std::vector<uint64_t> ranks;
    std::vector<HostHit> raw;
    pcagpu_stats_t stats;
};

namespace {

constexpr uint32_t kGroupMax = 1u << 20;   // scan groups per launch

// Test order on the device.  Tests whose inputs look random (about half the
// bits set, far from 0 and 0xFFFF) come first: few candidates map such an
// input to the target by accident, while inputs like 0 or 1 are matched by
// many constant-ish candidates.  Equal-target tests are interleaved
// round-robin so a near-constant candidate fails within two tests.  The
// order only decides when a lane stops testing, never which programs pass.
int fill_devtask(pcagpu *h)
{
    std::memset(&h->dt, 0, sizeof(h->dt));
    h->dt.T = h->T;
    h->dt.n_in = h->task.num_inputs;
    h->dt.max_cycles = (int)std::min<uint64_t>(h->task.max_cycles, 255);
    TestState roots[PCAGPU_MAX_TESTS];
    root_states(&h->task, roots);
    auto score = [&](int t) {
        int sc = 0;
        for (int i = 0; i < h->task.num_inputs; i++) {
            const uint32_t x = h->task.tests[t].in[i];
            const int pc = __builtin_popcount(x);
            sc += 64 - 8 * std::abs(pc - 8) + (int)(std::min<uint32_t>(x, 65536u - x) >> 10);
        }
        return sc;
    };
    std::map<uint16_t, std::vector<int>> by_val;
    for (int t = 0; t < h->T; t++) by_val[h->task.tests[t].out[0]].push_back(t);
    std::vector<std::vector<int>> groups;
    for (auto &kv : by_val) {
        std::vector<int> g = kv.second;
        std::stable_sort(g.begin(), g.end(), [&](int x, int y) { return score(x) > score(y); });
        groups.push_back(g);
    }
    std::stable_sort(groups.begin(), groups.end(), [&](const std::vector<int> &x, const std::vector<int> &y) {
        return score(x[0]) > score(y[0]);
    });
    std::vector<int> order;
    for (size_t r = 0; order.size() < (size_t)h->T; r++)
        for (auto &g : groups)
            if (r < g.size()) order.push_back(g[r]);
    for (int t = 0; t < h->T; t++) {
        h->dt.tgt[t] = h->task.tests[order[t]].out[0];
        h->dt.root[t] = roots[order[t]];
    }
    return 0;
}

int grid_for(pcagpu *h, uint64_t units)
{
    uint64_t waves = (uint64_t)h->cu_count * h->waves_per_cu;
    uint64_t wgs = (std::min<uint64_t>(units, waves) + SCAN_WAVES - 1) / SCAN_WAVES;
    return (int)std::max<uint64_t>(wgs, 1);
}

int run_scan(pcagpu *h, uint32_t n, uint32_t n_groups, uint64_t units)
{
    ScanArgs a;
    a.prefixes = h->d_batch.p;
    a.n_prefix = n;
    a.groups = h->d_groups.p;
    a.group_off = h->d_goff.p;
    a.n_groups = n_groups;
    a.n_units = units;
    a.D = h->D;
    a.s = h->s;
    a.k = h->k;
    a.cand = CandDev{h->d_lists.p, h->d_count.p, h->d_split.p};
    a.hits = h->d_hits.p;
    a.hit_cap = h->hit_cap;
    a.hit_count = h->d_hit_count.p;
    a.counters = h->d_counters.p;
    a.work = h->d_work.p;
    if (n_groups) {
        HIP_TRY(hipMemcpyAsync(h->d_groups.p, h->h_groups, (size_t)n_groups * sizeof(ScanGroup),
                               hipMemcpyHostToDevice, h->stream));
        HIP_TRY(hipMemcpyAsync(h->d_goff.p, h->h_goff, ((size_t)n_groups + 1) * sizeof(uint64_t),
                               hipMemcpyHostToDevice, h->stream));
    }
    HIP_TRY(hipMemsetAsync(h->d_work.p, 0, sizeof(unsigned long long), h->stream));
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    launch_scan(h->dt, a, grid_for(h, units), h->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->stats.kernel_ns += (uint64_t)((double)ms * 1e6);
    h->stats.launches++;
    return 0;
}

// All scan launches of one flush: k <= 1 in one launch; k >= 2 in launches
// of at most kGroupMax groups.
int scan_batch(pcagpu *h, uint32_t n)
{
    if (h->k <= 1) return run_scan(h, n, 0, n);
    const int mf_s = h->D - h->s - 1, mf_s1 = h->D - h->s - 2;
    uint32_t g = 0;
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t live = h->h_batch[i].live;
        const int ns = h->ct.count(live, mf_s);
        if (h->k == 2) {
            h->h_groups[g] = ScanGroup{i, 0, 0};
            h->h_goff[g] = acc;
            acc += (uint64_t)ns;
            g++;
        } else {
            for (int js = 0; js < ns; js++) {
                const uint32_t c = h->ct.at(live, mf_s, js);
                h->h_groups[g] = ScanGroup{i, (uint16_t)js, 0};
                h->h_goff[g] = acc;
                acc += (uint64_t)h->ct.count(live_after(live, c), mf_s1);
                g++;
            }
        }
        if (g + (uint32_t)CAND_LMAX + 64 > kGroupMax || i + 1 == n) {
            h->h_goff[g] = acc;
            int rc = run_scan(h, n, g, acc);
            if (rc) return rc;
            g = 0;
            acc = 0;
        }
    }
    return 0;
}

int read_counters(pcagpu *h, unsigned long long out[4])
{
    HIP_TRY(hipMemcpy(out, h->d_counters.p, 4 * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    return 0;
}

}  // namespace

extern "C" {

int pcagpu_task_load(const char *path, pca_task_t *task) { return task_load(path, task); }
int pcagpu_task_parse(const char *json, pca_task_t *task) { return task_parse(json, task); }
int pcagpu_task_evaluate(const pca_task_t *task, const uint16_t *prog, int len) { return task_evaluate(task, prog, len); }
void pcagpu_disasm(uint16_t insn, char *buf, size_t len) { disasm(insn, buf, len); }
int pcagpu_gen_candidates(uint16_t *out, int max, uint8_t live, int pos, int kernel_depth)
{
    return gen_candidates(out, max, live, pos, kernel_depth);
}

int pcagpu_open(const pca_task_t *task, const pcagpu_opts_t *opts, pcagpu_t **out)
{
    if (!task || !out) return PCAGPU_EINVAL;
    *out = nullptr;
    // gpu/hybrid.cu:648-659: single output, <= 7 inputs; ports index an
    // 8-entry port file
    if (task->num_outputs != 1 || task->num_inputs < 0 || task->num_inputs > 7 || task->num_tests < 0 ||
        task->num_tests > PCAGPU_MAX_TESTS)
        return PCAGPU_EINVAL;
    for (int i = 0; i < task->num_inputs; i++)
        if (task->input_ports[i] > 7) return PCAGPU_EINVAL;
    if (task->output_ports[0] > 7) return PCAGPU_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PCAGPU_ENODEV;
    pcagpu *h = new (std::nothrow) pcagpu();
    if (!h) return PCAGPU_ENOMEM;
    h->task = *task;
    h->T = task->num_tests;
    std::memset(&h->opts, 0, sizeof(h->opts));
    if (opts) h->opts = *opts;
    h->device = h->opts.device;
    if (h->device < 0 || h->device >= ndev) { delete h; return PCAGPU_EINVAL; }
    h->batch_cap = h->opts.batch_capacity ? h->opts.batch_capacity : (1u << 20);
    h->hit_cap = h->opts.hit_capacity ? h->opts.hit_capacity : (1u << 20);
    if (h->opts.waves_per_cu) h->waves_per_cu = h->opts.waves_per_cu;
    std::memset(&h->stats, 0, sizeof(h->stats));
    int rc = 0;
    auto fail = [&](int code) { pcagpu_close(h); return code; };
    if (hipSetDevice(h->device) != hipSuccess) return fail(PCAGPU_EHIP);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) return fail(PCAGPU_EHIP);
    h->devname = prop.name;
    h->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(PCAGPU_EHIP);
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) return fail(PCAGPU_EHIP);
    h->ct.build();
    fill_devtask(h);
    if ((rc = h->d_lists.ensure(h->ct.list.size())) || (rc = h->d_count.ensure(128)) ||
        (rc = h->d_split.ensure(128)) || (rc = h->d_batch.ensure(h->batch_cap)) ||
        (rc = h->d_groups.ensure(kGroupMax)) || (rc = h->d_goff.ensure(kGroupMax + 1)) ||
        (rc = h->d_hits.ensure(h->hit_cap)) ||
        (rc = h->d_hit_count.ensure(1)) || (rc = h->d_counters.ensure(8)) || (rc = h->d_work.ensure(1)))
        return fail(rc);
    if (hipHostMalloc((void **)&h->h_batch, (size_t)h->batch_cap * sizeof(pca_prefix_t)) != hipSuccess ||
        hipHostMalloc((void **)&h->h_groups, (size_t)kGroupMax * sizeof(ScanGroup)) != hipSuccess ||
        hipHostMalloc((void **)&h->h_goff, ((size_t)kGroupMax + 1) * sizeof(uint64_t)) != hipSuccess)
        return fail(PCAGPU_ENOMEM);
    if (hipMemcpy(h->d_lists.p, h->ct.list.data(), h->ct.list.size() * 2, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->d_count.p, h->ct.n, sizeof(h->ct.n), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(h->d_split.p, h->ct.a, sizeof(h->ct.a), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(h->d_counters.p, 0, 8 * sizeof(unsigned long long)) != hipSuccess)
        return fail(PCAGPU_EHIP);
    *out = h;
    return 0;
}

void pcagpu_close(pcagpu_t *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->d_lists.release();
    h->d_count.release();
    h->d_split.release();
    h->d_batch.release();
    h->d_groups.release();
    h->d_goff.release();
    h->d_hits.release();
    h->d_hit_count.release();
    h->d_counters.release();
    h->d_work.release();
    h->d_front[0].release();
    h->d_front[1].release();
    h->d_child_off.release();
    h->d_keep_cnt.release();
    h->d_scan_tmp.release();
    h->d_table.release();
    h->d_keep_bits.release();
    if (h->h_batch) (void)hipHostFree(h->h_batch);
    if (h->h_groups) (void)hipHostFree(h->h_groups);
    if (h->h_goff) (void)hipHostFree(h->h_goff);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char *pcagpu_device_name(pcagpu_t *h) { return h ? h->devname.c_str() : ""; }

int pcagpu_begin_depth(pcagpu_t *h, int target_depth, int split_depth)
{
    if (!h) return PCAGPU_EINVAL;
    if (target_depth < 0 || target_depth > PCAGPU_MAX_KDEPTH || split_depth < 0 || split_depth > target_depth)
        return PCAGPU_EINVAL;
    if (h->task.num_inputs + target_depth + 2 > PCAGPU_MAX_PROG) return PCAGPU_EINVAL;
    h->D = target_depth;
    h->s = split_depth;
    h->k = target_depth - split_depth;
    h->batch_n = 0;
    h->batch_first = 0;
    h->prefixes.clear();
    h->ranks.clear();
    h->raw.clear();
    h->in_depth = true;
    return 0;
}

int pcagpu_flush(pcagpu_t *h)
{
    if (!h || !h->in_depth) return PCAGPU_ESTATE;
    if (h->batch_n == 0) return (int)std::min<size_t>(h->raw.size(), 0x7FFFFFFF);
    (void)hipSetDevice(h->device);
    const uint32_t n = h->batch_n;
    HIP_TRY(hipMemcpyAsync(h->d_batch.p, h->h_batch, (size_t)n * sizeof(pca_prefix_t), hipMemcpyHostToDevice,
                           h->stream));
    uint32_t hits = 0;
    int rc;
    for (;;) {
        HIP_TRY(hipMemsetAsync(h->d_hit_count.p, 0, sizeof(uint32_t), h->stream));
        if ((rc = scan_batch(h, n))) return rc;
        HIP_TRY(hipMemcpy(&hits, h->d_hit_count.p, sizeof(uint32_t), hipMemcpyDeviceToHost));
        if (hits <= h->hit_cap) break;
        // grow and redo this batch: no hit may be dropped
        uint32_t want = hits + hits / 2 + 1024;
        h->d_hits.release();
        if ((rc = h->d_hits.ensure(want))) return rc;
        h->hit_cap = want;
        // the counters of the discarded pass must not be counted twice
        HIP_TRY(hipMemset(h->d_counters.p, 0, 4 * sizeof(unsigned long long)));
    }
    if (hits) {
        std::vector<HitRec> buf(hits);
        HIP_TRY(hipMemcpy(buf.data(), h->d_hits.p, (size_t)hits * sizeof(HitRec), hipMemcpyDeviceToHost));
        for (const HitRec &r : buf) {
            HostHit x;
            x.prefix = h->batch_first + r.prefix;
            x.mask = r.mask;
            x.flags = r.flags;
            std::memcpy(x.j, r.j, sizeof(x.j));
            h->raw.push_back(x);
        }
    }
    unsigned long long c[4];
    if ((rc = read_counters(h, c))) return rc;
    h->stats.candidates += c[0];
    h->stats.suffixes += c[1];
    h->stats.tests_evaluated += c[2];
    h->stats.parents += c[3];
    HIP_TRY(hipMemset(h->d_counters.p, 0, 4 * sizeof(unsigned long long)));
    h->stats.raw_hits += hits;
    h->batch_first += n;
    h->batch_n = 0;
    return (int)std::min<size_t>(h->raw.size(), 0x7FFFFFFF);
}

int pcagpu_submit(pcagpu_t *h, const pca_prefix_t *p, uint32_t n, uint64_t first_rank)
{
    if (!h || !h->in_depth) return PCAGPU_ESTATE;
    if (n && !p) return PCAGPU_EINVAL;
    for (uint32_t i = 0; i < n; i++) {
        if (p[i].num_insns != h->s) return PCAGPU_EINVAL;
        if (!h->ranks.empty() && first_rank + i <= h->ranks.back()) return PCAGPU_EINVAL;
        h->h_batch[h->batch_n++] = p[i];
        h->prefixes.push_back(p[i]);
        h->ranks.push_back(first_rank + i);
        h->stats.prefixes++;
        if (h->batch_n == h->batch_cap) {
            int rc = pcagpu_flush(h);
            if (rc < 0) return rc;
        }
    }
    return 0;
}

}  // extern "C"

namespace {

struct Node {
    uint64_t prefix;                 // depth-prefix index (rank order)
    uint16_t j[PCAGPU_MAX_KDEPTH];
    uint16_t insns[PCAGPU_MAX_KDEPTH];
    uint32_t live;                   // live mask after the full kernel
    uint8_t mask, dead;
};

bool node_less(const Node &a, const Node &b, int k)
{
    if (a.prefix != b.prefix) return a.prefix < b.prefix;
    for (int q = 0; q < k; q++)
        if (a.j[q] != b.j[q]) return a.j[q] < b.j[q];
    return false;
}

}  // namespace

extern "C" long pcagpu_end_depth(pcagpu_t *h, pca_solution_t *out, long cap)
{
    if (!h || !h->in_depth) return PCAGPU_ESTATE;
    int rc = pcagpu_flush(h);
    if (rc < 0) return rc;
    h->in_depth = false;
    const int D = h->D, s = h->s, k = h->k, T = h->T;
    // rebuild each hit's kernel from its prefix and candidate-list indices
    std::vector<Node> nodes;
    nodes.reserve(h->raw.size());
    for (const HostHit &x : h->raw) {
        Node nd;
        std::memset(&nd, 0, sizeof(nd));
        nd.prefix = x.prefix;
        const pca_prefix_t &p = h->prefixes[x.prefix];
        for (int q = 0; q < s; q++) nd.insns[q] = p.insns[q];
        uint32_t live = p.live;
        for (int q = 0; q < k; q++) {
            const int pos = s + q;
            const uint16_t c = h->ct.at(live, D - pos - 1, x.j[q]);
            nd.insns[pos] = c;
            nd.j[q] = x.j[q];
            live = live_after(live, c);
        }
        nd.live = live;
        nd.mask = x.mask;
        nd.dead = x.flags & 1;
        nodes.push_back(nd);
    }
    std::sort(nodes.begin(), nodes.end(), [k](const Node &a, const Node &b) { return node_less(a, b, k); });
    // Exact OEP at the last level (tools/enumerate.c:435-456, exact keys):
    // every node whose state has a target-matching register is a hit, so the
    // first visitor of each such state is among the hits.  Keep it, drop the
    // rest.  (k == 0: the prefixes were already deduplicated.)
    std::vector<char> keep(nodes.size(), 1);
    if (k >= 1) {
        ExactKeySet set(T * 17 > 0 ? T * 17 : 1);
        std::vector<TestState> root((size_t)(T > 0 ? T : 1)), st((size_t)(T > 0 ? T : 1));
        root_states(&h->task, root.data());
        std::vector<uint8_t> key((size_t)T * 17 + 1);
        for (size_t i = 0; i < nodes.size(); i++) {
            for (int t = 0; t < T; t++) {
                st[t] = root[t];
                for (int q = 0; q < D; q++) step_test(st[t], nodes[i].insns[q], (uint32_t)q);
            }
            oep_key_from_states(st.data(), T, key.data());
            keep[i] = (char)set.insert(key.data());
        }
    }
    long total = 0;
    for (size_t i = 0; i < nodes.size(); i++) {
        if (!keep[i] || nodes[i].dead) continue;
        const uint32_t solving = nodes[i].mask & nodes[i].live & 0xFEu;
        for (int r = 1; r < 8; r++) {
            if (!(solving >> r & 1u)) continue;
            if (out && total < cap) {
                pca_solution_t &so = out[total];
                std::memset(&so, 0, sizeof(so));
                so.prefix_rank = h->ranks[nodes[i].prefix];
                int n = 0;
                for (int q = 0; q < h->task.num_inputs; q++)
                    so.prog[n++] = enc_i(OP_IN, (uint32_t)q + 1, h->task.input_ports[q]);
                for (int q = 0; q < D; q++) so.prog[n++] = nodes[i].insns[q];
                so.prog[n++] = enc_i(OP_OUT, (uint32_t)r, h->task.output_ports[0]);
                so.prog[n++] = enc_r(OP_HLT, 0, 0, 0);
                so.len = (uint8_t)n;
                so.out_reg = (uint8_t)r;
                so.kernel_len = (uint8_t)D;
                // candidate path: prefix positions are recovered from the
                // grammar, suffix positions are the device indices
                uint32_t live = 0;
                for (int q = 0; q < h->task.num_inputs; q++) live |= 1u << (q + 1);
                for (int q = 0; q < D; q++) {
                    uint16_t cands[CAND_LMAX + 64];
                    int nc = gen_candidates(cands, CAND_LMAX + 64, live, q, D);
                    int idx = 0;
                    while (idx < nc && cands[idx] != nodes[i].insns[q]) idx++;
                    so.cand[q] = (uint16_t)idx;
                    live = live_after(live, nodes[i].insns[q]);
                }
                so.cpu_verified = task_evaluate(&h->task, so.prog, n) == h->task.num_tests ? 1 : 0;
            }
            total++;
        }
    }
    return total;
}

extern "C" long pcagpu_host_frontier(const pca_task_t *task, int D, int s, pca_prefix_t *out, long cap,
                                     uint64_t *nodes_out, uint64_t *pruned_out)
{
    if (!task || D < 0 || D > PCAGPU_MAX_KDEPTH || s < 0 || s > D) return PCAGPU_EINVAL;
    if (task->num_tests < 0 || task->num_tests > PCAGPU_MAX_TESTS) return PCAGPU_EINVAL;
    long n = 0;
    DfsStats st;
    dfs_prefixes(task, D, s,
                 [&](const pca_prefix_t &p) {
                     if (out && n < cap) out[n] = p;
                     n++;
                     // with an output buffer, stop once it is full: the first
                     // cap prefixes of the DFS order are what the caller wants
                     return !(out && n >= cap);
                 },
                 &st);
    if (nodes_out) *nodes_out = st.nodes;
    if (pruned_out) *pruned_out = st.pruned;
    return n;
}

extern "C" long pcagpu_raw_hits(pcagpu_t *h, uint16_t *rows, long cap)
{
    if (!h || !h->in_depth) return PCAGPU_ESTATE;
    std::vector<HostHit> v = h->raw;
    const int k = h->k;
    std::sort(v.begin(), v.end(), [k](const HostHit &a, const HostHit &b) {
        if (a.prefix != b.prefix) return a.prefix < b.prefix;
        for (int q = 0; q < k; q++)
            if (a.j[q] != b.j[q]) return a.j[q] < b.j[q];
        return false;
    });
    for (long i = 0; i < (long)v.size() && i < cap && rows; i++) {
        uint16_t *r = rows + i * 14;
        std::memset(r, 0, 14 * sizeof(uint16_t));
        for (int b = 0; b < 4; b++) r[b] = (uint16_t)(v[i].prefix >> (16 * b));
        r[4] = v[i].flags;
        r[5] = v[i].mask;
        for (int q = 0; q < k; q++) r[6 + q] = v[i].j[q];
    }
    return (long)v.size();
}

extern "C" int pcagpu_get_stats(pcagpu_t *h, pcagpu_stats_t *st)
{
    if (!h || !st) return PCAGPU_EINVAL;
    *st = h->stats;
    return 0;
}

extern "C" void pcagpu_reset_stats(pcagpu_t *h)
{
    if (h) std::memset(&h->stats, 0, sizeof(h->stats));
}

// ------------------------------------------------------------ wavefront
extern "C" long pcagpu_frontier(pcagpu_t *h, int D, int s, pca_prefix_t *out, long cap)
{
    if (!h) return PCAGPU_EINVAL;
    if (D < 0 || D > PCAGPU_MAX_KDEPTH || s < 0 || s > D) return PCAGPU_EINVAL;
    (void)hipSetDevice(h->device);
    int rc;
    // level 0: the root
    pca_prefix_t root;
    std::memset(&root, 0, sizeof(root));
    for (int i = 0; i < h->task.num_inputs; i++) root.live |= (uint8_t)(1u << (i + 1));
    if ((rc = h->d_front[0].ensure(1))) return rc;
    HIP_TRY(hipMemcpy(h->d_front[0].p, &root, sizeof(root), hipMemcpyHostToDevice));
    uint64_t n = 1;
    int cur = 0;
    for (int pos = 0; pos < s; pos++) {
        ExpandArgs a;
        std::memset(&a, 0, sizeof(a));
        a.parents = h->d_front[cur].p;
        a.n_parents = n;
        a.D = D;
        a.pos = pos;
        a.cand = CandDev{h->d_lists.p, h->d_count.p, h->d_split.p};
        if ((rc = h->d_child_off.ensure(n + 1)) || (rc = h->d_keep_cnt.ensure(n + 1)) ||
            (rc = h->d_scan_tmp.ensure(exclusive_scan_tmp_words(n) + 16)))
            return rc;
        HIP_TRY(hipMemsetAsync(h->d_child_off.p, 0, (n + 1) * sizeof(uint64_t), h->stream));
        HIP_TRY(hipEventRecord(h->ev0, h->stream));
        launch_expand_count(a, h->d_child_off.p, h->stream);
        launch_exclusive_scan(h->d_child_off.p, n, h->d_scan_tmp.p, h->stream);
        uint64_t children = 0;
        HIP_TRY(hipMemcpyAsync(&children, h->d_child_off.p + n, sizeof(uint64_t), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        uint64_t tcap = 1024;
        while (tcap < 2 * children) tcap <<= 1;
        if ((rc = h->d_table.ensure(2 * tcap)) || (rc = h->d_keep_bits.ensure(children / 32 + 8)))
            return rc;
        // keys 0 (empty), min slots all-ones so atomicMin keeps the smallest index
        launch_table_init(h->d_table.p, tcap, h->stream);
        HIP_TRY(hipMemsetAsync(h->d_keep_bits.p, 0, (children / 32 + 8) * sizeof(uint32_t), h->stream));
        HIP_TRY(hipMemsetAsync(h->d_counters.p + 4, 0, 2 * sizeof(unsigned long long), h->stream));
        a.child_off = h->d_child_off.p;
        a.table = h->d_table.p;
        a.table_mask = tcap - 1;
        a.keep_bits = h->d_keep_bits.p;
        a.keep_count = h->d_keep_cnt.p;
        a.counters = h->d_counters.p + 4;
        a.work = h->d_work.p;
        const int grid = grid_for(h, n);
        HIP_TRY(hipMemsetAsync(h->d_work.p, 0, sizeof(unsigned long long), h->stream));
        launch_expand_insert(h->dt, a, grid, h->stream);
        HIP_TRY(hipMemsetAsync(h->d_work.p, 0, sizeof(unsigned long long), h->stream));
        HIP_TRY(hipMemsetAsync(h->d_keep_cnt.p, 0, (n + 1) * sizeof(uint64_t), h->stream));
        launch_expand_survive(h->dt, a, grid, h->stream);
        launch_exclusive_scan(h->d_keep_cnt.p, n, h->d_scan_tmp.p, h->stream);
        uint64_t kept = 0;
        unsigned long long ctr[2] = {0, 0};
        HIP_TRY(hipMemcpyAsync(&kept, h->d_keep_cnt.p + n, sizeof(uint64_t), hipMemcpyDeviceToHost, h->stream));
        HIP_TRY(hipMemcpyAsync(ctr, h->d_counters.p + 4, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                               h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        if (ctr[0] || ctr[1]) {
            h->stats.oep_collisions += ctr[0];
            return ctr[0] ? PCAGPU_ECOLLISION : PCAGPU_EOVERFLOW;
        }
        const int nxt = cur ^ 1;
        if ((rc = h->d_front[nxt].ensure(kept ? kept : 1))) return rc;
        a.keep_off = h->d_keep_cnt.p;
        a.out = h->d_front[nxt].p;
        HIP_TRY(hipMemsetAsync(h->d_work.p, 0, sizeof(unsigned long long), h->stream));
        launch_expand_compact(a, grid, h->stream);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipEventRecord(h->ev1, h->stream));
        HIP_TRY(hipStreamSynchronize(h->stream));
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->stats.expand_ns += (uint64_t)((double)ms * 1e6);
        h->stats.frontier_nodes += children;
        h->stats.frontier_kept += kept;
        n = kept;
        cur = nxt;
        if (n == 0) break;
    }
    if (out && n) {
        const uint64_t m = std::min<uint64_t>(n, (uint64_t)std::max<long>(cap, 0));
        if (m) HIP_TRY(hipMemcpy(out, h->d_front[cur].p, m * sizeof(pca_prefix_t), hipMemcpyDeviceToHost));
    }
    return (long)n;
}

// ------------------------------------------------------------ IDDFS
extern "C" int pcagpu_solve(pcagpu_t *h, const pcagpu_solve_opts_t *o, pcagpu_result_t *res,
                            pca_solution_t *out, long cap)
{
    if (!h || !o || !res) return PCAGPU_EINVAL;
    std::memset(res, 0, sizeof(*res));
    res->optimal_depth = -1;
    res->depths_done = -1;
    const double t0 = now_s();
    const int maxd = std::min(o->max_depth, PCAGPU_MAX_PROG - 2 - h->task.num_inputs);
    for (int D = std::max(o->start_depth, 0); D <= maxd && D <= PCAGPU_MAX_KDEPTH; D++) {
        int s = o->split < 0 ? (D > 0 ? D - 1 : 0) : std::min(o->split, D);
        int rc = pcagpu_begin_depth(h, D, s);
        if (rc) return rc;
        const uint64_t cand0 = h->stats.candidates, pre0 = h->stats.prefixes;
        bool stopped = false;
        if (o->producer == 0) {
            long n = pcagpu_frontier(h, D, s, nullptr, 0);
            if (n < 0) return (int)n;
            std::vector<pca_prefix_t> f((size_t)n);
            if (n) {
                long m = pcagpu_frontier(h, D, s, f.data(), n);
                if (m < 0) return (int)m;
            }
            for (long i = 0; i < n; i += h->batch_cap) {
                uint32_t c = (uint32_t)std::min<long>(h->batch_cap, n - i);
                if ((rc = pcagpu_submit(h, f.data() + i, c, (uint64_t)i))) return rc;
                if (o->time_limit > 0 && now_s() - t0 > o->time_limit) { stopped = true; break; }
            }
        } else {
            uint64_t rank = 0;
            DfsStats ds;
            int err = 0;
            dfs_prefixes(&h->task, D, s,
                         [&](const pca_prefix_t &p) {
                             int r = pcagpu_submit(h, &p, 1, rank++);
                             if (r) { err = r; return false; }
                             if (o->time_limit > 0 && (rank & 4095) == 0 && now_s() - t0 > o->time_limit) {
                                 stopped = true;
                                 return false;
                             }
                             return true;
                         },
                         &ds);
            if (err) return err;
        }
        long nsol = pcagpu_end_depth(h, out, cap);
        if (nsol < 0) return (int)nsol;
        res->candidates[D] = h->stats.candidates - cand0;
        res->prefixes[D] = h->stats.prefixes - pre0;
        if (o->verbose)
            std::printf("  depth %d (split %d): %llu prefixes, %llu candidates, %ld solutions, %.3f s\n", D, s,
                        (unsigned long long)res->prefixes[D], (unsigned long long)res->candidates[D], nsol,
                        now_s() - t0);
        if (stopped) { res->timed_out = 1; break; }
        res->depths_done = D;
        if (nsol > 0) {
            res->optimal_depth = D;
            res->num_solutions = nsol;
            break;
        }
    }
    res->seconds = now_s() - t0;
    return 0;
}
