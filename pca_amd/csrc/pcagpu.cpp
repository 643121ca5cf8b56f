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
#include <memory>
#include <mutex>
#include <set>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "host_dfs.h"
#include "pcagpu_impl.h"

using namespace pca;

namespace pca {

static_assert(NBKINDS == NBK && VGMAX == VG_MAX && LMAX == CAND_LMAX, "host and device grammar tables differ");

void set_test_order(pcagpu *h, const std::vector<int> &order)
{
    for (int t = 0; t < h->T; t++) {
        h->dt.tgt[t] = h->task.tests[order[t]].out[0];
        h->dt.tgtw[t] = h->dt.tgt[t];
        for (int i = 0; i < 12; i++)
            if (kLdiImms[i] == h->dt.tgt[t]) h->dt.tgtw[t] |= 1u << (16 + i);
        h->dt.root[t] = h->roots[order[t]];
    }
    h->order = order;
}

void devtask_for(const pca_task_t *task, DevTask *dt)
{
    std::memset(dt, 0, sizeof(*dt));
    dt->T = task->num_tests;
    dt->n_in = task->num_inputs;
    dt->max_cycles = (int)std::min<uint64_t>(task->max_cycles, 255);
    TestState roots[PCAGPU_MAX_TESTS];
    root_states(task, roots);
    const std::vector<int> order = device_test_order(task);
    for (int t = 0; t < task->num_tests; t++) {
        dt->tgt[t] = task->tests[order[t]].out[0];
        dt->tgtw[t] = dt->tgt[t];
        for (int i = 0; i < 12; i++)
            if (kLdiImms[i] == dt->tgt[t]) dt->tgtw[t] |= 1u << (16 + i);
        dt->root[t] = roots[order[t]];
    }
}

int fill_devtask(pcagpu *h)
{
    std::memset(&h->dt, 0, sizeof(h->dt));
    h->dt.T = h->T;
    h->dt.n_in = h->task.num_inputs;
    h->dt.max_cycles = (int)std::min<uint64_t>(h->task.max_cycles, 255);
    root_states(&h->task, h->roots);
    std::memset(h->t_reach, 0, sizeof(h->t_reach));
    std::memset(h->t_end, 0, sizeof(h->t_end));
    set_test_order(h, device_test_order(&h->task));
    return 0;
}

// Adaptive test order.  A launch reports, per device position, how many
// candidates still alive after the first two tests reached it and how many
// it ended; accumulated per task test, tests 2.. are re-sorted by that end
// rate, highest first (a test not reached yet counts 1/2), so that the
// tests which actually separate this task's near-misses run early.  The
// first two positions stay: the per-op masks run there for every group.
// The order decides only when a candidate stops being tested, never which
// candidates pass.
int adapt_test_order(pcagpu *h, const unsigned long long *ts)
{
    for (int t = 0; t < h->T; t++) {
        h->t_reach[h->order[t]] += ts[2 * t];
        h->t_end[h->order[t]] += ts[2 * t + 1];
    }
    if (h->T <= 3) return 0;
    std::vector<int> ord = h->order;
    auto rate = [&](int x) { return h->t_reach[x] ? (double)h->t_end[x] / (double)h->t_reach[x] : 0.5; };
    std::stable_sort(ord.begin() + 2, ord.end(), [&](int x, int y) { return rate(x) > rate(y); });
    if (ord == h->order) return 0;
    set_test_order(h, ord);
    HIP_TRY(hipMemcpy(h->d_task.p, &h->dt, sizeof(DevTask), hipMemcpyHostToDevice));
    return 0;
}

void grammar_image(bool branchless, GrammarDev *img)
{
    const CandTable &ct = branchless ? shared_grammar_branchless() : shared_grammar();
    std::memset(img, 0, sizeof(GrammarDev));
    for (int m = 0; m < 128; m++) {
        std::memcpy(img->lists[m], &ct.list[(size_t)m * CAND_LMAX], sizeof(img->lists[m]));
        img->count[m] = ct.n[m];
        img->split[m] = ct.a[m];
        img->cand_total[m] = ct.cand_total[m];
        std::memcpy(img->vg_rdmask[m], &ct.vg_rdmask[(size_t)m * VG_MAX], sizeof(img->vg_rdmask[m]));
        std::memcpy(img->vg_j[m], &ct.vg_j[(size_t)m * VG_MAX * 8], sizeof(img->vg_j[m]));
        std::memcpy(img->bits_init[m], &ct.bits_init[(size_t)m * NBK], sizeof(img->bits_init[m]));
        std::memcpy(img->bits_gidx[m], &ct.bits_gidx[(size_t)m * NBK * 64], sizeof(img->bits_gidx[m]));
        for (int l = 0; l < 64; l++) {
            uint32_t kinds = 0;
            for (int q = 0; q < NBKINDS; q++) kinds |= (uint32_t)((img->bits_init[m][q] >> l) & 1ull) << q;
            img->kind_bits[m][l] = (uint16_t)kinds;
        }
    }
}

const GrammarDev *shared_grammar_dev(int device, bool branchless)
{
    static std::mutex mu;
    static std::map<int, const GrammarDev *> tables;
    std::lock_guard<std::mutex> lock(mu);
    const int key = 2 * device + (branchless ? 1 : 0);
    auto it = tables.find(key);
    if (it != tables.end()) return it->second;
    std::unique_ptr<GrammarDev> img(new (std::nothrow) GrammarDev);
    if (!img) return nullptr;
    grammar_image(branchless, img.get());
    GrammarDev *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(GrammarDev)) != hipSuccess) return nullptr;
    if (hipMemcpy(d, img.get(), sizeof(GrammarDev), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(d);
        return nullptr;
    }
    tables[key] = d;
    return d;
}

CandDev cand_dev(const GrammarDev *g)
{
    return CandDev{g, g->count, g->split, g->cand_total, &g->bits_init[0][0]};
}

CandDev cand_dev(const pcagpu *h) { return cand_dev(h->d_grammar); }

int grid_for(pcagpu *h, uint64_t units, bool multi, uint32_t levels)
{
    uint32_t wpc = h->waves_per_cu;
    if (!wpc) {
        // persistent grid: as many resident workgroups as the scan kernel's
        // registers and LDS (which grows with the state levels) allow
        const uint32_t lv = levels ? std::min<uint32_t>(levels, (uint32_t)MAXLEV) : 3u;
        uint32_t &c = h->scan_wpc[multi ? 1 : 0][lv];
        if (!c) {
            int b = 0;
            c = (scan_occupancy(multi, lv, &b) == hipSuccess && b > 0) ? (uint32_t)b * SCAN_WAVES : 16u;
        }
        wpc = c;
    }
    uint64_t waves = (uint64_t)h->cu_count * wpc;
    uint64_t wgs = (std::min<uint64_t>(units, waves) + SCAN_WAVES - 1) / SCAN_WAVES;
    return (int)std::max<uint64_t>(wgs, 1);
}

int grid_for_aux(pcagpu *h, uint64_t units, AuxKernel which)
{
    uint32_t wpc = h->waves_per_cu;
    if (!wpc) {
        uint32_t &c = h->aux_wpc[(int)which];
        if (!c) {
            int b = 0;
            c = (aux_occupancy(which, &b) == hipSuccess && b > 0) ? (uint32_t)b * SCAN_WAVES : 16u;
        }
        wpc = c;
    }
    const uint64_t waves = (uint64_t)h->cu_count * wpc;
    const uint64_t wgs = (std::min<uint64_t>(units, waves) + SCAN_WAVES - 1) / SCAN_WAVES;
    return (int)std::max<uint64_t>(wgs, 1);
}

// d_counters: [0..3] a scan launch's counters, [4..5] the wavefront's, [6]
// the ALU digest's, [XCNT..XCNT+1] the scan's shared-scan counters.
constexpr int XCNT = 8;

// Scan tile autotuner.  The tile (children per block, tests run by wave
// masks) is chosen per suffix length k: the first launches of that shape
// each try one candidate on real work, and the one with the least kernel
// time per unit of work (group tests and parent states, weighted as in the
// roofline) is kept for the rest of the handle's life.
// A launch with too few parents to time does not count as a try.
// (blocks of 128 children are cut to 64 where only 64 fit: per-child
// scans and suffixes of other lengths)
constexpr ScanTile kTiles[] = {{64, 2, TILE_SHARED}, {64, 1, TILE_SHARED}, {64, 3, TILE_SHARED},
                               {128, 3, TILE_SHARED}, {128, 2, TILE_SHARED}};
constexpr int kNTiles = (int)(sizeof(kTiles) / sizeof(kTiles[0]));
static_assert(kNTiles <= 8, "pcagpu::tile_cost holds 8 tries");
constexpr uint64_t kTileMinParents = 200000;

ScanTile current_tile(const pcagpu *h, int *cand)
{
    *cand = -1;
    // bits 16..23 of the option: 1 scans every child on its own
    const uint8_t mode = ((h->opts.scan_tile >> 16) & 0xFFu) == 1u ? TILE_PER_CHILD : TILE_SHARED;
    // blocks above 64 children only where the shared last position scans them
    const uint32_t bk_max = (mode == TILE_SHARED && (h->k == 2 || h->k == 3)) ? BLOCK_KIDS_SHARED : BLOCK_KIDS_MAX;
    if (h->opts.scan_tile & 0xFFFFu) {
        const uint32_t bk = h->opts.scan_tile & 0xFFu, mt = (h->opts.scan_tile >> 8) & 0xFFu;
        return ScanTile{(uint16_t)std::min<uint32_t>(std::max<uint32_t>(bk, 1u), bk_max),
                        (uint8_t)std::min<uint32_t>(std::max<uint32_t>(mt, 1u), 6u), mode};
    }
    const int k = std::min(std::max(h->k, 0), (int)MAXLEV);
    ScanTile t;
    if (h->tile_best[k] >= 0) {
        t = kTiles[h->tile_best[k]];
    } else {
        *cand = h->tile_tried[k];
        t = kTiles[*cand];
    }
    t.mode = mode;
    t.block_kids = (uint16_t)std::min<uint32_t>(t.block_kids, bk_max);
    return t;
}

// (a launch queued before the previous one's feedback arrived may have been
// given the same candidate: only the candidate being tried is credited)
void tile_feedback(pcagpu *h, int cand, int k_launch, float ms, uint64_t parents, uint64_t tests)
{
    const int k = std::min(std::max(k_launch, 0), (int)MAXLEV);
    if (cand < 0 || cand != h->tile_tried[k] || parents < kTileMinParents) return;
    // kernel time per unit of the launch's work (9 ops per group test
    // deciding a last-position value group, 15 per parent and test): the
    // tries run on different batches, whose work per parent differs
    const double work = 9.0 * (double)tests + 15.0 * (double)h->dt.T * (double)parents;
    h->tile_cost[k][cand] = (double)ms * 1e6 / work;
    if (++h->tile_tried[k] < kNTiles) return;
    int best = 0;
    for (int i = 1; i < kNTiles; i++)
        if (h->tile_cost[k][i] < h->tile_cost[k][best]) best = i;
    h->tile_best[k] = best;
}

// Read-back area of one scan launch (pinned host memory): counters,
// hit count, per-test survivor counts.
struct LaunchOut {
    // [0..3] the scan counters, [XCNT..XCNT+1] its shared-scan counters
    // (the words between belong to other kernels and are read back unused)
    unsigned long long cnt[XCNT + 2];
    unsigned long long ts[2 * MAXT];
    uint32_t hits;
    uint32_t pad;
};

// One scan launch through the handle's graph (built on the first call):
// prefix and block upload, an event marking the uploads done, clears of the
// launch counters, the kernel between two event records, read-back of the
// counters.  Per launch only the upload sizes and the kernel parameters
// change.  Queued on the handle's stream; nothing waits here.
int graph_scan(pcagpu *h, const ScanLaunch &l, const ScanItem &it, int grid, uint32_t n, uint32_t n_blocks)
{
    ScanNodeArgs ka;
    hipKernelNodeParams kp;
    scan_node_params(l, it, h->dt, grid, &ka, &kp);
    const size_t pbytes = (size_t)std::max<uint32_t>(n, 1u) * sizeof(pca_prefix_t);
    const size_t bbytes = (size_t)std::max<uint32_t>(n_blocks, 1u) * sizeof(ScanBlock);
    LaunchOut *o = reinterpret_cast<LaunchOut *>(h->h_out);
    if (!h->g_exec) {
        hipGraphNode_t zero, up, ev0, ev1, r0, r1, r2;
        uint32_t *zp_ptr[4] = {reinterpret_cast<uint32_t *>(h->d_work.p), reinterpret_cast<uint32_t *>(h->d_tstat.p),
                               reinterpret_cast<uint32_t *>(h->d_counters.p), h->d_hit_count.p};
        const uint32_t zp_n[4] = {2, 4 * MAXT, 2 * (XCNT + 2), 1};
        ZeroNodeArgs za;
        hipKernelNodeParams zp;
        zero_node_params(zp_ptr, zp_n, &za, &zp);
        HIP_TRY(hipGraphCreate(&h->g_graph, 0));
        HIP_TRY(hipGraphAddMemcpyNode1D(&h->g_prefixes, h->g_graph, nullptr, 0, h->d_batch.p, h->h_batch, pbytes,
                                        hipMemcpyHostToDevice));
        HIP_TRY(hipGraphAddMemcpyNode1D(&h->g_blocks, h->g_graph, &h->g_prefixes, 1, h->d_blocks.p, h->h_blocks,
                                        bbytes, hipMemcpyHostToDevice));
        HIP_TRY(hipGraphAddEventRecordNode(&up, h->g_graph, &h->g_blocks, 1, h->ev_up));
        HIP_TRY(hipGraphAddKernelNode(&zero, h->g_graph, &up, 1, &zp));
        HIP_TRY(hipGraphAddEventRecordNode(&ev0, h->g_graph, &zero, 1, h->ev0));
        HIP_TRY(hipGraphAddKernelNode(&h->g_kernel, h->g_graph, &ev0, 1, &kp));
        HIP_TRY(hipGraphAddEventRecordNode(&ev1, h->g_graph, &h->g_kernel, 1, h->ev1));
        HIP_TRY(hipGraphAddMemcpyNode1D(&r0, h->g_graph, &ev1, 1, o->cnt, h->d_counters.p, sizeof(o->cnt),
                                        hipMemcpyDeviceToHost));
        HIP_TRY(hipGraphAddMemcpyNode1D(&r1, h->g_graph, &r0, 1, o->ts, h->d_tstat.p, sizeof(o->ts),
                                        hipMemcpyDeviceToHost));
        HIP_TRY(hipGraphAddMemcpyNode1D(&r2, h->g_graph, &r1, 1, &o->hits, h->d_hit_count.p, sizeof(o->hits),
                                        hipMemcpyDeviceToHost));
        HIP_TRY(hipGraphInstantiate(&h->g_exec, h->g_graph, nullptr, nullptr, 0));
    } else {
        HIP_TRY(hipGraphExecMemcpyNodeSetParams1D(h->g_exec, h->g_prefixes, h->d_batch.p, h->h_batch, pbytes,
                                                  hipMemcpyHostToDevice));
        HIP_TRY(hipGraphExecMemcpyNodeSetParams1D(h->g_exec, h->g_blocks, h->d_blocks.p, h->h_blocks, bbytes,
                                                  hipMemcpyHostToDevice));
        HIP_TRY(hipGraphExecKernelNodeSetParams(h->g_exec, h->g_kernel, &kp));
    }
    HIP_TRY(hipGraphLaunch(h->g_exec, h->stream));
    return 0;
}

// The same launch with plain stream calls.  Uploads of the staged prefixes
// and blocks are optional: prefixes already on the device (a device-resident
// depth) are read in place, and the redo of a launch whose hits overflowed
// reads what the first run uploaded.
int eager_scan(pcagpu *h, const ScanLaunch &l, const ScanItem &it, int grid, uint32_t n, uint32_t n_blocks,
               bool up_prefixes, bool up_blocks)
{
    LaunchOut *o = reinterpret_cast<LaunchOut *>(h->h_out);
    if (up_prefixes)
        HIP_TRY(hipMemcpyAsync(h->d_batch.p, h->h_batch, (size_t)n * sizeof(pca_prefix_t), hipMemcpyHostToDevice,
                               h->stream));
    if (up_blocks && n_blocks)
        HIP_TRY(hipMemcpyAsync(h->d_blocks.p, h->h_blocks, (size_t)n_blocks * sizeof(ScanBlock),
                               hipMemcpyHostToDevice, h->stream));
    if (up_prefixes || up_blocks) HIP_TRY(hipEventRecord(h->ev_up, h->stream));
    HIP_TRY(hipMemsetAsync(h->d_work.p, 0, sizeof(unsigned long long), h->stream));
    HIP_TRY(hipMemsetAsync(h->d_tstat.p, 0, 2 * MAXT * sizeof(unsigned long long), h->stream));
    HIP_TRY(hipMemsetAsync(h->d_counters.p, 0, (XCNT + 2) * sizeof(unsigned long long), h->stream));
    HIP_TRY(hipMemsetAsync(h->d_hit_count.p, 0, sizeof(uint32_t), h->stream));
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    launch_scan(l, &it, h->dt, grid, h->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipMemcpyAsync(o->cnt, h->d_counters.p, sizeof(o->cnt), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(o->ts, h->d_tstat.p, sizeof(o->ts), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipMemcpyAsync(&o->hits, h->d_hit_count.p, sizeof(o->hits), hipMemcpyDeviceToHost, h->stream));
    return 0;
}

// The kernel arguments of a one-item scan launch over the device batch.
int scan_args(pcagpu *h, const pca_prefix_t *src, uint32_t n, uint64_t units, ScanTile tile, int k, ScanItem *it,
              ScanLaunch *l)
{
    std::memset(it, 0, sizeof(*it));
    it->task = h->d_task.p;
    it->prefixes = src;
    it->blocks = h->d_blocks.p;
    it->n_prefix = n;
    it->unit_begin = 0;
    it->n_units = units;
    it->D = h->D;
    it->s = h->D - k;
    it->k = k;
    it->tile = tile;
    it->hits = h->d_hits.p;
    it->hit_count = h->d_hit_count.p;
    it->hit_cap = h->hit_cap;
    l->items = h->d_items.p;
    l->n_items = 1;
    l->levels = scan_levels(k);
    l->n_units = units;
    l->cand = cand_dev(h);
    l->counters = h->d_counters.p;
    l->work = h->d_work.p;
    l->tstat = h->opts.fixed_test_order ? nullptr : h->d_tstat.p;
    const int grid = grid_for(h, units, false, l->levels);
    l->grab = scan_grab(k, units, grid);
    l->pad = 0;
    l->xcounters = h->d_counters.p + XCNT;
    return grid;
}

// Queues one scan launch over n prefixes and h_blocks[0, n_blocks) (the
// previous one must have been collected): the staged h_batch (src ==
// d_batch, uploaded by the launch) or prefixes already on the device at src.
// Returns at once; the host may refill the staging buffers after
// wait_uploads.
int start_scan(pcagpu *h, const pca_prefix_t *src, uint32_t n, uint32_t n_blocks, uint64_t units, ScanTile tile,
               int cand)
{
    ScanItem it;
    ScanLaunch l;
    const int grid = scan_args(h, src, n, units, tile, h->k, &it, &l);
    const bool staged = src == h->d_batch.p;
    int rc = !staged ? eager_scan(h, l, it, grid, n, n_blocks, false, true)
             : h->opts.eager_flush ? eager_scan(h, l, it, grid, n, n_blocks, true, true)
                                   : graph_scan(h, l, it, grid, n, n_blocks);
    if (rc) return rc;
    h->pend.on = true;
    h->pend.src = src;
    h->pend.n = n;
    h->pend.n_blocks = n_blocks;
    h->pend.units = units;
    h->pend.tile = tile;
    h->pend.cand = cand;
    h->pend.k = h->k;
    h->pend.batch_first = h->batch_first;
    h->up_waited = false;
    return 0;
}

// Blocks until the launch in flight has read its prefixes and blocks from
// the pinned staging buffers (usually long done: the uploads come first).
int wait_uploads(pcagpu *h)
{
    if (h->up_waited) return 0;
    if (h->pend.on) HIP_TRY(hipEventSynchronize(h->ev_up));
    h->up_waited = true;
    return 0;
}

int settle_scan(pcagpu *h)
{
    if (!h->pend.on) return 0;
    h->pend.on = false;
    h->up_waited = true;
    const LaunchOut *o = reinterpret_cast<const LaunchOut *>(h->h_out);
    const auto p = h->pend;
    HIP_TRY(hipStreamSynchronize(h->stream));
    float ms = 0.f;
    for (;;) {
        HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->stats.kernel_ns += (uint64_t)((double)ms * 1e6);
        h->stats.launches++;
        if (o->hits <= h->hit_cap) break;
        // overflow: a larger buffer, and the kernel again over the batch
        // still on the device
        const uint32_t want = o->hits + o->hits / 2 + 1024;
        h->d_hits.release();
        h->hit_cap = 0;   // until the larger buffer exists: launches write no records
        int rc = h->d_hits.ensure(want);
        if (rc) return rc;
        h->hit_cap = want;
        ScanItem it;
        ScanLaunch l;
        const int grid = scan_args(h, p.src, p.n, p.units, p.tile, p.k, &it, &l);
        if ((rc = eager_scan(h, l, it, grid, p.n, p.n_blocks, false, false))) return rc;
        HIP_TRY(hipStreamSynchronize(h->stream));
    }
    h->stats.scan_tile =
        (uint64_t)p.tile.block_kids | ((uint64_t)p.tile.mask_tests << 8) | ((uint64_t)p.tile.mode << 16);
    h->stats.candidates += o->cnt[0];
    h->stats.suffixes += o->cnt[1];
    h->stats.tests_evaluated += o->cnt[2];
    h->stats.tests_executed += o->cnt[2] - o->cnt[XCNT] + o->cnt[XCNT + 1];
    h->stats.parents += o->cnt[3];
    const uint32_t hits = o->hits;
    h->stats.raw_hits += hits;
    if (hits) {
        std::vector<HitRec> buf(hits);
        HIP_TRY(hipMemcpy(buf.data(), h->d_hits.p, (size_t)hits * sizeof(HitRec), hipMemcpyDeviceToHost));
        for (const HitRec &r : buf) {
            HostHit x;
            x.prefix = p.batch_first + r.prefix;
            x.mask = r.mask;
            x.flags = r.flags;
            std::memcpy(x.j, r.j, sizeof(x.j));
            h->raw.push_back(x);
        }
    }
    tile_feedback(h, p.cand, p.k, ms, o->cnt[3], o->cnt[2]);
    if (h->opts.fixed_test_order) return 0;
    unsigned long long ts[2 * MAXT];
    std::memcpy(ts, o->ts, sizeof(ts));
    return adapt_test_order(h, ts);
}

bool prefix_blocks(const CandTable &ct, int D, int s, uint32_t live, uint32_t i, ScanBlock *out, uint32_t cap,
                   uint32_t *nb, int bk)
{
    const int mf_s = D - s - 1, mf_s1 = D - s - 2;
    const int ns = ct.count(live, mf_s);
    uint32_t b = *nb;
    auto push = [&](uint32_t js, int nc) {
        for (int j0 = 0; j0 < nc; j0 += bk) {
            if (b >= cap) return false;
            out[b++] = ScanBlock{i, (uint16_t)js, (uint16_t)j0};
        }
        return true;
    };
    if (D - s == 2) {
        if (!push(0, ns)) return false;
    } else {
        for (int js = 0; js < ns; js++) {
            const uint32_t c = ct.at(live, mf_s, js);
            if (!push((uint32_t)js, ct.count(live_after(live, c), mf_s1))) return false;
        }
    }
    *nb = b;
    return true;
}


// Block buffers (host and device) at full size on the first scan with
// k >= 2; handles that only scan single suffix positions never need them.
int ensure_block_buffers(pcagpu *h)
{
    if (h->block_cap >= kBlockMax) return 0;
    int rc;
    if ((rc = h->d_blocks.ensure(kBlockMax))) return rc;
    if (h->h_blocks) (void)hipHostFree(h->h_blocks);
    h->h_blocks = nullptr;
    h->block_cap = 0;
    if (hipHostMalloc((void **)&h->h_blocks, (size_t)kBlockMax * sizeof(ScanBlock)) != hipSuccess)
        return PCAGPU_ENOMEM;
    h->block_cap = kBlockMax;
    return 0;
}

// Queues the scan launches of one flush: k <= 1 in one launch (a unit per
// prefix); k >= 2 in launches of at most kBlockMax blocks.  Each launch
// first collects the one before it (they share the hit buffer and the
// counters); the last one is left in flight.  For k >= 2 the blocks are
// built while the previous launch still runs.
// hostp: the batch's prefixes on the host (their live masks make the
// blocks); src: the same on the device (d_batch: uploaded from h_batch by
// the launch).  Hits are numbered from h->batch_first.
int scan_batch(pcagpu *h, const pca_prefix_t *hostp, const pca_prefix_t *src, uint32_t n)
{
    int cand, rc;
    ScanTile tile;
    if (h->k <= 1) {
        if ((rc = settle_scan(h))) return rc;
        tile = current_tile(h, &cand);
        return start_scan(h, src, n, 0, n, tile, cand);
    }
    if ((rc = wait_uploads(h)) || (rc = ensure_block_buffers(h))) return rc;
    tile = current_tile(h, &cand);
    uint32_t nb = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (prefix_blocks(h->ct, h->D, h->s, hostp[i].live, i, h->h_blocks, h->block_cap, &nb, tile.block_kids))
            continue;
        if (nb == 0) return PCAGPU_EOVERFLOW;
        if ((rc = settle_scan(h)) || (rc = start_scan(h, src, n, nb, nb, tile, cand)) || (rc = wait_uploads(h)))
            return rc;
        if (h->deadline > 0 && now_s() > h->deadline) {
            h->stopped = true;   // the time limit passed: the rest of the batch is not scanned
            return 0;
        }
        tile = current_tile(h, &cand);
        nb = 0;
        i--;
    }
    if (!nb) return 0;
    if ((rc = settle_scan(h))) return rc;
    return start_scan(h, src, n, nb, nb, tile, cand);
}

size_t handle_device_bytes(const pcagpu *h)
{
    return h->d_batch.bytes() + h->d_blocks.bytes() + h->d_groups.bytes() + h->d_goff.bytes() + h->d_task.bytes() + h->d_items.bytes() +
           h->d_hits.bytes() + h->d_hit_count.bytes() + h->d_counters.bytes() + h->d_work.bytes() + h->d_tstat.bytes() +
           h->d_front[0].bytes() + h->d_front[1].bytes() + h->d_child_off.bytes() + h->d_keep_cnt.bytes() +
           h->d_scan_tmp.bytes() + h->d_table.bytes() + h->d_keep_bits.bytes() + h->d_child_hash.bytes() + h->d_parent_rows.bytes() + h->d_q0.bytes() + h->d_qf.bytes() +
           h->d_match.bytes() + h->d_match_count.bytes();
}

int stage_prefixes(pcagpu *h, const pca_prefix_t *p, uint32_t n, uint64_t first_rank)
{
    if (!h->in_depth) return PCAGPU_ESTATE;
    int rc = wait_uploads(h);
    if (rc) return rc;
    if (h->batch_n + n > h->batch_cap) return PCAGPU_EOVERFLOW;
    for (uint32_t i = 0; i < n; i++) {
        if (p[i].num_insns != h->s) return PCAGPU_EINVAL;
        if (!h->ranks.empty() && first_rank + i <= h->ranks.back()) return PCAGPU_EINVAL;
        h->h_batch[h->batch_n++] = p[i];
        h->prefixes.push_back(p[i]);
        h->ranks.push_back(first_rank + i);
        h->stats.prefixes++;
    }
    return 0;
}

uint32_t prefixes_fitting(const pcagpu *h, const pca_prefix_t *p, uint32_t n, ScanTile tile)
{
    if (h->k <= 1) return n;
    const int mf_s = h->D - h->s - 1, mf_s1 = h->D - h->s - 2;
    const int bk = tile.block_kids;
    auto blocks = [bk](int nc) { return (uint64_t)((nc + bk - 1) / bk); };
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t live = p[i].live;
        const int ns = h->ct.count(live, mf_s);
        uint64_t b = 0;
        if (h->k == 2) {
            b = blocks(ns);
        } else {
            for (int js = 0; js < ns; js++)
                b += blocks(h->ct.count(live_after(live, h->ct.at(live, mf_s, js)), mf_s1));
        }
        if (total + b > kBlockMax) return i;
        total += b;
    }
    return n;
}

int build_blocks(pcagpu *h, ScanTile tile, uint32_t *n_blocks, uint64_t *n_units)
{
    return build_blocks_of(h, h->h_batch, h->batch_n, tile, n_blocks, n_units);
}

int build_blocks_of(pcagpu *h, const pca_prefix_t *hostp, uint32_t n, ScanTile tile, uint32_t *n_blocks,
                    uint64_t *n_units)
{
    *n_blocks = 0;
    if (h->k <= 1) {
        *n_units = n;
        return 0;
    }
    int rc;
    if ((rc = wait_uploads(h)) || (rc = ensure_block_buffers(h))) return rc;
    uint32_t nb = 0;
    for (uint32_t i = 0; i < n; i++)
        if (!prefix_blocks(h->ct, h->D, h->s, hostp[i].live, i, h->h_blocks, h->block_cap, &nb, tile.block_kids))
            return PCAGPU_EOVERFLOW;
    *n_blocks = nb;
    *n_units = nb;
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

int pcagpu_state_records(const pca_task_t *task, const uint16_t *insns, const uint8_t *len, int n, uint16_t *rec,
                         uint8_t *live)
{
    if (!task || n < 0 || (n && (!insns || !len || !rec || !live))) return PCAGPU_EINVAL;
    if (task->num_tests < 0 || task->num_tests > PCAGPU_MAX_TESTS || task->num_inputs < 0 || task->num_inputs > 7)
        return PCAGPU_EINVAL;
    TestState roots[PCAGPU_MAX_TESTS];
    root_states(task, roots);
    uint32_t live0 = 0;
    for (int i = 0; i < task->num_inputs; i++) live0 |= 1u << (i + 1);
    for (int i = 0; i < n; i++) {
        const uint16_t *p = insns + (size_t)i * PCAGPU_MAX_KDEPTH;
        const int L = len[i];
        if (L > PCAGPU_MAX_KDEPTH) return PCAGPU_EINVAL;
        uint32_t lv = live0;
        for (int q = 0; q < L; q++) lv = live_after(lv, p[q]);
        live[i] = (uint8_t)lv;
        uint16_t *r = rec + (size_t)i * 32 * 16;
        std::memset(r, 0, 32 * 16 * sizeof(uint16_t));
        for (int t = 0; t < task->num_tests; t++) {
            TestState s = roots[t];
            for (int q = 0; q < L; q++) step_test(s, p[q], (uint32_t)q);
            uint16_t *w = r + t * 16;
            w[0] = task->num_inputs > 0 ? task->tests[t].in[0] : 0;
            w[1] = task->num_inputs > 1 ? task->tests[t].in[1] : 0;
            w[2] = task->tests[t].out[0];
            for (int k = 0; k < 8; k++) w[3 + k] = s.r[k];
            w[11] = s.flags;
            w[12] = s.frozen;
            w[13] = s.resume;
        }
    }
    return 0;
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
    h->waves_per_cu = h->opts.waves_per_cu;
    std::memset(&h->stats, 0, sizeof(h->stats));
    int rc = 0;
    auto fail = [&](int code) { pcagpu_close(h); return code; };
    if (hipSetDevice(h->device) != hipSuccess) return fail(PCAGPU_EHIP);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) != hipSuccess) return fail(PCAGPU_EHIP);
    h->devname = prop.name[0] ? prop.name : prop.gcnArchName;
    h->cu_count = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) return fail(PCAGPU_EHIP);
    if (hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&h->ev_up, hipEventDisableTiming) != hipSuccess)
        return fail(PCAGPU_EHIP);
    fill_devtask(h);
    if (!(h->d_grammar = shared_grammar_dev(h->device))) return fail(PCAGPU_ENOMEM);
    if ((rc = h->d_batch.ensure(h->batch_cap)) ||
        (rc = h->d_blocks.ensure(1)) ||
        (rc = h->d_hits.ensure(h->hit_cap)) ||
        (rc = h->d_hit_count.ensure(1)) || (rc = h->d_counters.ensure(XCNT + 2)) || (rc = h->d_work.ensure(1)) ||
        (rc = h->d_tstat.ensure(2 * MAXT)))
        return fail(rc);
    if (hipHostMalloc((void **)&h->h_batch, (size_t)h->batch_cap * sizeof(pca_prefix_t)) != hipSuccess ||
        hipHostMalloc((void **)&h->h_blocks, sizeof(ScanBlock)) != hipSuccess ||
        hipHostMalloc(&h->h_out, sizeof(LaunchOut)) != hipSuccess)
        return fail(PCAGPU_ENOMEM);
    if (hipMemset(h->d_counters.p, 0, (XCNT + 2) * sizeof(unsigned long long)) != hipSuccess)
        return fail(PCAGPU_EHIP);
    if ((rc = h->d_task.ensure(1)) || (rc = h->d_items.ensure(64))) return fail(rc);
    if (hipMemcpy(h->d_task.p, &h->dt, sizeof(DevTask), hipMemcpyHostToDevice) != hipSuccess)
        return fail(PCAGPU_EHIP);
    *out = h;
    return 0;
}

int pcagpu_retask(pcagpu_t *h, const pca_task_t *task)
{
    if (!h || !task) return PCAGPU_EINVAL;
    if (task->num_outputs != 1 || task->num_inputs < 0 || task->num_inputs > 7 || task->num_tests < 0 ||
        task->num_tests > PCAGPU_MAX_TESTS || task->output_ports[0] > 7)
        return PCAGPU_EINVAL;
    for (int i = 0; i < task->num_inputs; i++)
        if (task->input_ports[i] > 7) return PCAGPU_EINVAL;
    (void)hipSetDevice(h->device);
    int rc = settle_scan(h);
    if (rc) return rc;
    h->task = *task;
    h->T = task->num_tests;
    fill_devtask(h);
    HIP_TRY(hipMemcpy(h->d_task.p, &h->dt, sizeof(DevTask), hipMemcpyHostToDevice));
    h->in_depth = false;
    h->batch_n = 0;
    return 0;
}

void pcagpu_close(pcagpu_t *h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->d_batch.release();
    h->d_blocks.release();
    h->d_groups.release();
    h->d_goff.release();
    h->d_task.release();
    h->d_items.release();
    h->d_hits.release();
    h->d_hit_count.release();
    h->d_counters.release();
    h->d_tstat.release();
    h->d_work.release();
    h->d_front[0].release();
    h->d_front[1].release();
    h->d_child_off.release();
    h->d_keep_cnt.release();
    h->d_scan_tmp.release();
    h->d_table.release();
    h->d_keep_bits.release();
    h->d_child_hash.release();
    h->d_parent_rows.release();
    h->d_q0.release();
    h->d_qf.release();
    h->d_match.release();
    h->d_match_count.release();
    if (h->h_batch) (void)hipHostFree(h->h_batch);
    if (h->h_front) (void)hipHostFree(h->h_front);
    if (h->g_exec) (void)hipGraphExecDestroy(h->g_exec);
    if (h->g_graph) (void)hipGraphDestroy(h->g_graph);
    if (h->h_out) (void)hipHostFree(h->h_out);
    if (h->h_blocks) (void)hipHostFree(h->h_blocks);
    if (h->h_groups) (void)hipHostFree(h->h_groups);
    if (h->h_goff) (void)hipHostFree(h->h_goff);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->ev_up) (void)hipEventDestroy(h->ev_up);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char *pcagpu_device_name(pcagpu_t *h) { return h ? h->devname.c_str() : ""; }

int pcagpu_begin_depth(pcagpu_t *h, int target_depth, int split_depth)
{
    if (!h) return PCAGPU_EINVAL;
    int rc = settle_scan(h);
    if (rc) return rc;
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
    h->dev_front = false;
    h->d_src = nullptr;
    h->n_front = 0;
    h->in_depth = true;
    return 0;
}

int pcagpu_flush(pcagpu_t *h)
{
    if (!h || !h->in_depth) return PCAGPU_ESTATE;
    (void)hipSetDevice(h->device);
    int rc;
    if (h->batch_n) {
        // the staged batch, or a range of a device-resident depth
        const uint32_t n = h->batch_n;
        const pca_prefix_t *hp = h->dev_front ? h->h_front + h->batch_first : h->h_batch;
        const pca_prefix_t *src = h->dev_front ? h->d_src + h->batch_first : h->d_batch.p;
        if ((rc = scan_batch(h, hp, src, n))) return rc;
        h->batch_first += n;
        h->batch_n = 0;
    }
    if ((rc = settle_scan(h))) return rc;
    return (int)std::min<size_t>(h->raw.size(), 0x7FFFFFFF);
}

int pcagpu_submit(pcagpu_t *h, const pca_prefix_t *p, uint32_t n, uint64_t first_rank)
{
    if (!h || !h->in_depth || h->dev_front) return PCAGPU_ESTATE;
    if (n && !p) return PCAGPU_EINVAL;
    for (uint32_t i = 0; i < n; i++) {
        if (p[i].num_insns != h->s) return PCAGPU_EINVAL;
        if (!h->ranks.empty() && first_rank + i <= h->ranks.back()) return PCAGPU_EINVAL;
        int rc;
        if (!h->up_waited && (rc = wait_uploads(h))) return rc;
        h->h_batch[h->batch_n++] = p[i];
        h->prefixes.push_back(p[i]);
        h->ranks.push_back(first_rank + i);
        h->stats.prefixes++;
        if (h->batch_n == h->batch_cap) {
            // a full batch goes to the device and runs while the caller
            // produces the next one; the next full batch (or a flush)
            // collects it
            (void)hipSetDevice(h->device);
            const uint32_t nb = h->batch_n;
            if ((rc = scan_batch(h, h->h_batch, h->d_batch.p, nb))) return rc;
            h->batch_first += nb;
            h->batch_n = 0;
        }
    }
    return 0;
}

}  // extern "C"

namespace {

// ---- node paths: a node below a depth prefix is (prefix index, candidate
// indices at positions s, s+1, ...); lexicographic order = DFS rank order.
struct Path {
    uint32_t prefix = 0;
    uint8_t len = 0;
    uint16_t j[PCAGPU_MAX_KDEPTH] = {0, 0, 0, 0, 0, 0, 0, 0};
    bool operator<(const Path &o) const
    {
        if (prefix != o.prefix) return prefix < o.prefix;
        const int n = len < o.len ? len : o.len;
        for (int q = 0; q < n; q++)
            if (j[q] != o.j[q]) return j[q] < o.j[q];
        return len < o.len;
    }
    bool operator==(const Path &o) const
    {
        if (prefix != o.prefix || len != o.len) return false;
        for (int q = 0; q < len; q++)
            if (j[q] != o.j[q]) return false;
        return true;
    }
};

Path parent_of(const Path &x)
{
    Path p = x;
    p.len = (uint8_t)(x.len - 1);
    p.j[p.len] = 0;
    return p;
}

// kernel instructions of a node (positions 0 .. s+len-1) and its live mask
int node_insns(const pcagpu *h, const Path &x, uint16_t *insns, uint32_t *live_out)
{
    const pca_prefix_t &p = h->pfx()[x.prefix];
    for (int q = 0; q < h->s; q++) insns[q] = p.insns[q];
    uint32_t live = p.live;
    for (int q = 0; q < x.len; q++) {
        const int pos = h->s + q;
        const uint16_t c = h->ct.at(live, h->D - pos - 1, x.j[q]);
        insns[pos] = c;
        live = live_after(live, c);
    }
    if (live_out) *live_out = live;
    return h->s + x.len;
}

// exact OEP key of a node: per test (device order) r0..r7 and flags
void node_key(const pcagpu *h, const Path &x, uint8_t *key, TestState *st)
{
    uint16_t insns[PCAGPU_MAX_KDEPTH];
    const int n = node_insns(h, x, insns, nullptr);
    for (int t = 0; t < h->T; t++) {
        st[t] = h->dt.root[t];
        for (int q = 0; q < n; q++) step_test(st[t], insns[q], (uint32_t)q);
    }
    oep_key_from_states(st, h->T, key);
}

uint64_t h_mix64(uint64_t x)
{
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return x;
}

// host twin of the device row hash (kernels.hip row_words_hash)
uint64_t h_row_hash(uint64_t h, const TestState &s)
{
    const uint64_t w0 = (uint64_t)s.r[1] | ((uint64_t)s.r[2] << 16) | ((uint64_t)s.r[3] << 32) | ((uint64_t)s.r[4] << 48);
    const uint64_t w1 = (uint64_t)s.r[5] | ((uint64_t)s.r[6] << 16) | ((uint64_t)s.r[7] << 32) | ((uint64_t)s.flags << 48);
    h = h_mix64(h ^ w0);
    return h_mix64(h ^ w1 ^ 0x5851F42D4C957F2Dull);
}

// Builds the scan groups of prefixes [first, first+n) for kk positions after
// the split and calls launch(n_groups, n_units) for every filled chunk.
template <typename F>
int for_group_chunks(pcagpu *h, uint32_t first, uint32_t n, int kk, F &&launch)
{
    if (kk <= 1) return launch(0u, (uint64_t)n);
    const int mf_s = h->D - h->s - 1, mf_s1 = h->D - h->s - 2;
    uint32_t g = 0;
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t live = h->pfx()[first + i].live;
        const int ns = h->ct.count(live, mf_s);
        if (kk == 2) {
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
            int rc = launch(g, acc);
            if (rc) return rc;
            g = 0;
            acc = 0;
        }
    }
    return 0;
}

// Device pass: every node with kk positions after the split (all depth
// prefixes) whose state hashes like one of the queries.  Returns candidate
// matches (path, query id); the caller confirms them exactly.
int run_match(pcagpu *h, int kk, const std::vector<uint64_t> &qh0, const std::vector<uint64_t> &qhf,
              std::vector<std::pair<Path, uint32_t>> &out)
{
    int rc;
    uint64_t cap0 = 64, capf = 64;
    while (cap0 < 2 * qh0.size()) cap0 <<= 1;
    while (capf < 2 * qhf.size()) capf <<= 1;
    std::vector<unsigned long long> t0(cap0, 0ull), tf(2 * capf, 0ull);
    for (uint64_t k : qh0) {
        uint64_t sl = (k >> 1) & (cap0 - 1);
        while (t0[sl] != 0ull && t0[sl] != k) sl = (sl + 1) & (cap0 - 1);
        t0[sl] = k;
    }
    for (size_t q = 0; q < qhf.size(); q++) {
        const uint64_t k = qhf[q];
        uint64_t sl = (k >> 1) & (capf - 1);
        while (tf[2 * sl] != 0ull && tf[2 * sl] != k) sl = (sl + 1) & (capf - 1);
        if (tf[2 * sl] == k) continue;   // equal full hashes: one id answers for both (host confirms)
        tf[2 * sl] = k;
        tf[2 * sl + 1] = q;
    }
    if ((rc = h->d_q0.ensure(cap0)) || (rc = h->d_qf.ensure(2 * capf)) || (rc = h->d_match_count.ensure(1)) ||
        (rc = h->d_groups.ensure(kGroupMax)) || (rc = h->d_goff.ensure(kGroupMax + 1)))
        return rc;
    if (!h->h_groups && hipHostMalloc((void **)&h->h_groups, (size_t)kGroupMax * sizeof(ScanGroup)) != hipSuccess)
        return PCAGPU_ENOMEM;
    if (!h->h_goff && hipHostMalloc((void **)&h->h_goff, ((size_t)kGroupMax + 1) * sizeof(uint64_t)) != hipSuccess)
        return PCAGPU_ENOMEM;
    if (h->d_match.n == 0 && (rc = h->d_match.ensure(1 << 16))) return rc;
    HIP_TRY(hipMemcpy(h->d_q0.p, t0.data(), cap0 * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(h->d_qf.p, tf.data(), 2 * capf * 8, hipMemcpyHostToDevice));
    const uint32_t total = (uint32_t)h->n_pfx();
    for (uint32_t first = 0; first < total; first += h->batch_cap) {
        const uint32_t n = std::min<uint32_t>(h->batch_cap, total - first);
        // a device-resident depth is read in place
        const pca_prefix_t *dp = h->dev_front ? h->d_src + first : h->d_batch.p;
        if (!h->dev_front)
            HIP_TRY(hipMemcpy(h->d_batch.p, h->prefixes.data() + first, (size_t)n * sizeof(pca_prefix_t),
                              hipMemcpyHostToDevice));
        rc = for_group_chunks(h, first, n, kk, [&](uint32_t n_groups, uint64_t units) -> int {
            if (n_groups) {
                HIP_TRY(hipMemcpy(h->d_groups.p, h->h_groups, (size_t)n_groups * sizeof(ScanGroup),
                                  hipMemcpyHostToDevice));
                HIP_TRY(hipMemcpy(h->d_goff.p, h->h_goff, ((size_t)n_groups + 1) * 8, hipMemcpyHostToDevice));
            }
            for (;;) {
                MatchArgs a;
                a.prefixes = dp;
                a.n_prefix = n;
                a.groups = h->d_groups.p;
                a.group_off = h->d_goff.p;
                a.n_groups = n_groups;
                a.n_units = units;
                a.D = h->D;
                a.s = h->s;
                a.kk = kk;
                a.cand = cand_dev(h);
                a.q0 = h->d_q0.p;
                a.q0_mask = cap0 - 1;
                a.qf = h->d_qf.p;
                a.qf_mask = capf - 1;
                a.out = h->d_match.p;
                a.cap = (uint32_t)h->d_match.n;
                a.count = h->d_match_count.p;
                a.work = h->d_work.p;
                HIP_TRY(hipMemsetAsync(h->d_match_count.p, 0, 4, h->stream));
                HIP_TRY(hipMemsetAsync(h->d_work.p, 0, 8, h->stream));
                launch_match(h->dt, a, grid_for_aux(h, units, AuxKernel::Match), h->stream);
                HIP_TRY(hipGetLastError());
                uint32_t cnt = 0;
                HIP_TRY_S(h->stream, hipMemcpyAsync(&cnt, h->d_match_count.p, 4, hipMemcpyDeviceToHost, h->stream));
                HIP_TRY_S(h->stream, hipStreamSynchronize(h->stream));
                if (cnt > a.cap) {
                    h->d_match.release();
                    int r2 = h->d_match.ensure((size_t)cnt + cnt / 2 + 1024);
                    if (r2) return r2;
                    continue;
                }
                std::vector<MatchRec> buf(cnt);
                if (cnt)
                    HIP_TRY(hipMemcpy(buf.data(), h->d_match.p, (size_t)cnt * sizeof(MatchRec), hipMemcpyDeviceToHost));
                for (const MatchRec &r : buf) {
                    Path p;
                    p.prefix = first + r.prefix;
                    p.len = (uint8_t)kk;
                    for (int q = 0; q < kk; q++) p.j[q] = r.j[q];
                    out.push_back({p, r.query});
                }
                return 0;
            }
        });
        if (rc) return rc;
    }
    return 0;
}

struct Node {
    Path path;
    uint64_t skey;                   // (prefix, j0, j1) packed, the order of paths of length <= 2
    uint32_t live;                   // live mask after the full kernel
    uint8_t mask, dead;
};

// OEP keys (KL bytes each) of nodes[idx[a..b)] into keys; consecutive nodes
// below one prefix reuse that prefix's replayed per-test states.
void node_keys(const pcagpu *h, const std::vector<Node> &nodes, const std::vector<size_t> &idx, size_t a, size_t b,
               uint8_t *keys, int KL)
{
    TestState base[MAXT], st[MAXT];
    uint32_t cached = 0xFFFFFFFFu;
    for (size_t q = a; q < b; q++) {
        const Path &x = nodes[idx[q]].path;
        uint16_t insns[PCAGPU_MAX_KDEPTH];
        const int n = node_insns(h, x, insns, nullptr);
        if (x.prefix != cached) {
            for (int t = 0; t < h->T; t++) {
                base[t] = h->dt.root[t];
                for (int i = 0; i < h->s; i++) step_test(base[t], insns[i], (uint32_t)i);
            }
            cached = x.prefix;
        }
        for (int t = 0; t < h->T; t++) {
            st[t] = base[t];
            for (int i = h->s; i < n; i++) step_test(st[t], insns[i], (uint32_t)i);
        }
        oep_key_from_states(st, h->T, keys + (q - a) * (size_t)KL);
    }
}

// A hash of each node's OEP key (the same per-test rows node_keys writes,
// chained through h_row_hash), for nodes[idx[a..b)]: equal keys give equal
// hashes, so only nodes of equal hash need their keys compared.
void node_hashes(const pcagpu *h, const std::vector<Node> &nodes, const std::vector<size_t> &idx, size_t a,
                 size_t b, uint64_t *out)
{
    TestState base[MAXT], st[MAXT];
    uint32_t cached = 0xFFFFFFFFu;
    for (size_t q = a; q < b; q++) {
        const Path &x = nodes[idx[q]].path;
        uint16_t insns[PCAGPU_MAX_KDEPTH];
        const int n = node_insns(h, x, insns, nullptr);
        if (x.prefix != cached) {
            for (int t = 0; t < h->T; t++) {
                base[t] = h->dt.root[t];
                for (int i = 0; i < h->s; i++) step_test(base[t], insns[i], (uint32_t)i);
            }
            cached = x.prefix;
        }
        // a light multiply-xor chain over the rows' 20 bytes (a collision
        // only costs a byte-for-byte compare), mixed once at the end
        uint64_t hf = 0x243F6A8885A308D3ull;
        for (int t = 0; t < h->T; t++) {
            st[t] = base[t];
            for (int i = h->s; i < n; i++) step_test(st[t], insns[i], (uint32_t)i);
            uint64_t w[3];
            std::memcpy(w, &st[t], 16);
            w[2] = (uint64_t)st[t].flags;
            hf = (hf ^ w[0]) * 0x9E3779B97F4A7C15ull;
            hf = (hf ^ w[1]) * 0xC2B2AE3D27D4EB4Full;
            hf = (hf ^ w[2]) * 0x165667B19E3779F9ull;
        }
        out[q - a] = h_mix64(hf) | 1ull;
    }
}

// f(lo, hi) over [0, n) in contiguous runs on up to 8 host threads (one
// when n is small)
template <typename F>
void parallel_runs(size_t n, F &&f)
{
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t nt = n < 8192 ? 1 : std::min<size_t>(std::min(hw, 8u), (n + 4095) / 4096);
    if (nt <= 1) { f((size_t)0, n); return; }
    std::vector<std::thread> th;
    const size_t per = (n + nt - 1) / nt;
    for (size_t t = 1; t < nt; t++) {
        const size_t lo = t * per, hi = std::min(n, lo + per);
        if (lo < hi) th.emplace_back([&f, lo, hi] { f(lo, hi); });
    }
    f((size_t)0, std::min(n, per));
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" long pcagpu_end_depth(pcagpu_t *h, pca_solution_t *out, long cap)
{
    if (!h || !h->in_depth) return PCAGPU_ESTATE;
    int rc = pcagpu_flush(h);
    if (rc < 0) return rc;
    h->in_depth = false;
    const int D = h->D, k = h->k, T = h->T;
    const int KL = T * 17 > 0 ? T * 17 : 1;
    // Only a hit that can be output (no dead test, a live matching register)
    // needs a verdict, and only a node of the same state can prune it --
    // which has the same register-match mask.  Hits of other masks take no
    // further part (a task can have hundreds of thousands of them).
    bool mask_used[256] = {};
    std::vector<uint32_t> live_of(h->raw.size());
    parallel_runs(h->raw.size(), [&](size_t lo, size_t hi) {
        for (size_t i = lo; i < hi; i++) {
            const HostHit &x = h->raw[i];
            Path p;
            p.prefix = (uint32_t)x.prefix;
            p.len = (uint8_t)k;
            for (int q = 0; q < k; q++) p.j[q] = x.j[q];
            uint16_t insns[PCAGPU_MAX_KDEPTH];
            node_insns(h, p, insns, &live_of[i]);
        }
    });
    for (size_t i = 0; i < h->raw.size(); i++) {
        const HostHit &x = h->raw[i];
        if (!(x.flags & 1) && (x.mask & live_of[i] & 0xFEu)) mask_used[x.mask] = true;
    }
    std::vector<size_t> sel;
    for (size_t i = 0; i < h->raw.size(); i++)
        if (mask_used[h->raw[i].mask]) sel.push_back(i);
    std::vector<Node> nodes(sel.size());
    parallel_runs(sel.size(), [&](size_t lo, size_t hi) {
        for (size_t q = lo; q < hi; q++) {
            const size_t i = sel[q];
            const HostHit &x = h->raw[i];
            Node &nd = nodes[q];
            nd.path.prefix = (uint32_t)x.prefix;
            nd.path.len = (uint8_t)k;
            for (int p = 0; p < k; p++) nd.path.j[p] = x.j[p];
            nd.live = live_of[i];
            nd.mask = x.mask;
            nd.dead = x.flags & 1;
            nd.skey = ((uint64_t)nd.path.prefix << 24) | ((uint64_t)(k >= 1 ? nd.path.j[0] : 0) << 12) |
                      (uint64_t)(k >= 2 ? nd.path.j[1] : 0);
        }
    });
    // canonical (DFS rank) order; candidate indices are below 4096.  Runs
    // sorted on host threads, then merged pairwise.
    auto by_rank = [k](const Node &a, const Node &b) { return k <= 2 ? a.skey < b.skey : a.path < b.path; };
    {
        std::mutex mu;
        std::vector<std::pair<size_t, size_t>> runs;
        parallel_runs(nodes.size(), [&](size_t lo, size_t hi) {
            std::sort(nodes.begin() + lo, nodes.begin() + hi, by_rank);
            std::lock_guard<std::mutex> lk(mu);
            runs.emplace_back(lo, hi);
        });
        std::sort(runs.begin(), runs.end());
        while (runs.size() > 1) {   // merge neighbouring runs pairwise
            std::vector<std::pair<size_t, size_t>> next;
            for (size_t r = 0; r < runs.size(); r += 2) {
                if (r + 1 == runs.size()) {
                    next.push_back(runs[r]);
                    break;
                }
                std::inplace_merge(nodes.begin() + runs[r].first, nodes.begin() + runs[r].second,
                                   nodes.begin() + runs[r + 1].second, by_rank);
                next.emplace_back(runs[r].first, runs[r + 1].second);
            }
            runs.swap(next);
        }
    }
    std::vector<TestState> st((size_t)(T > 0 ? T : 1));
    std::vector<uint8_t> key((size_t)KL);

    // ---- exact OEP through the suffix levels (k >= 2).  A node at level p
    // survives iff its parent survives and it is a BR or no lower-ranked
    // node of the same state has a surviving parent (the reference's DFS
    // first visitor, tools/enumerate.c:435-456, with exact keys).  Needed
    // verdicts are collected top-down (parents of hits, then parents of the
    // lower-ranked same-state nodes the device finds) and resolved bottom-up.
    std::map<Path, char> survive;
    if (k >= 2) {
        std::vector<std::set<Path>> X((size_t)k);
        for (const Node &nd : nodes) X[k - 1].insert(parent_of(nd.path));
        std::map<Path, std::vector<Path>> earlier;
        std::map<Path, char> is_br;
        for (int L = k - 1; L >= 1; L--) {
            std::vector<Path> xs(X[L].begin(), X[L].end());
            std::map<std::vector<uint8_t>, uint32_t> qid;
            std::vector<std::vector<uint8_t>> qkeys;
            std::vector<uint64_t> qh0, qhf;
            std::vector<uint32_t> xq(xs.size(), 0xFFFFFFFFu);
            for (size_t i = 0; i < xs.size(); i++) {
                uint16_t insns[PCAGPU_MAX_KDEPTH];
                const int n = node_insns(h, xs[i], insns, nullptr);
                const bool br = insn_op(insns[n - 1]) == OP_BR;
                is_br[xs[i]] = br;
                if (L > 1) X[L - 1].insert(parent_of(xs[i]));
                if (br) continue;
                node_key(h, xs[i], key.data(), st.data());
                auto it = qid.find(key);
                if (it == qid.end()) {
                    const uint32_t id = (uint32_t)qkeys.size();
                    qid.emplace(key, id);
                    qkeys.push_back(key);
                    qh0.push_back(h_row_hash(0x243F6A8885A308D3ull, st[0]) | 1ull);
                    uint64_t hf = 0x243F6A8885A308D3ull;
                    for (int t = 0; t < T; t++) hf = h_row_hash(hf, st[t]);
                    qhf.push_back(hf | 1ull);
                    xq[i] = id;
                } else {
                    xq[i] = it->second;
                }
            }
            if (qkeys.empty()) continue;
            std::vector<std::pair<Path, uint32_t>> m;
            if ((rc = run_match(h, L, qh0, qhf, m))) return rc;
            // confirm on the host, byte for byte, and bucket by query
            std::vector<std::vector<Path>> byq(qkeys.size());
            std::vector<uint8_t> k2((size_t)KL);
            for (auto &pm : m) {
                node_key(h, pm.first, k2.data(), st.data());
                auto it = qid.find(k2);
                if (it != qid.end()) byq[it->second].push_back(pm.first);
                else h->stats.oep_collisions++;   // hash-equal, state-different: ignored
            }
            for (auto &v : byq) std::sort(v.begin(), v.end());
            for (size_t i = 0; i < xs.size(); i++) {
                if (xq[i] == 0xFFFFFFFFu) continue;
                std::vector<Path> &e = earlier[xs[i]];
                for (const Path &y : byq[xq[i]]) {
                    if (!(y < xs[i])) break;
                    e.push_back(y);
                    if (L > 1) X[L - 1].insert(parent_of(y));
                }
            }
        }
        for (int L = 1; L <= k - 1; L++) {
            for (const Path &x : X[L]) {
                const bool ps = L == 1 ? true : survive[parent_of(x)] != 0;
                if (!ps) { survive[x] = 0; continue; }
                if (is_br[x]) { survive[x] = 1; continue; }
                bool pruned = false;
                for (const Path &y : earlier[x])
                    if (L == 1 || survive[parent_of(y)]) { pruned = true; break; }
                survive[x] = pruned ? 0 : 1;
            }
        }
    }
    // ---- last level: among hit nodes with a surviving parent, the first of
    // each state (every node of a target-matching state is a hit).  Only a
    // node that can be output (no dead test, a live matching register) needs
    // the verdict, and only an earlier node of the same state can take it
    // away: every visited node's key is hashed on host threads; the nodes
    // whose hash equals an outputtable node's are then keyed in full, in
    // rank order per hash, and compared byte for byte.
    std::vector<char> keep(nodes.size(), 0);
    {
        std::vector<size_t> idx;   // visited nodes, rank order
        idx.reserve(nodes.size());
        for (size_t i = 0; i < nodes.size(); i++) {
            if (k >= 2 && !survive[parent_of(nodes[i].path)]) continue;
            if (k == 0) { keep[i] = 1; continue; }
            idx.push_back(i);
        }
        std::vector<uint64_t> hs(idx.size());
        parallel_runs(idx.size(), [&](size_t lo, size_t hi) { node_hashes(h, nodes, idx, lo, hi, hs.data() + lo); });
        auto outputtable = [&](size_t i) { return !nodes[i].dead && (nodes[i].mask & nodes[i].live & 0xFEu) != 0u; };
        // the outputtable nodes' hashes in a small open-addressed table; a
        // slot's list collects, in rank order, every visited node of that hash
        size_t cap = 16;
        size_t n_out = 0;
        for (size_t q = 0; q < idx.size(); q++) n_out += outputtable(idx[q]) ? 1 : 0;
        while (cap < 2 * n_out) cap <<= 1;
        std::vector<uint64_t> tab(cap, 0);
        auto slot_of = [&](uint64_t hv, bool add) -> long {
            for (size_t j = (size_t)(h_mix64(hv) & (cap - 1));; j = (j + 1) & (cap - 1)) {
                if (tab[j] == hv) return (long)j;
                if (tab[j] == 0) {
                    if (!add) return -1;
                    tab[j] = hv;
                    return (long)j;
                }
            }
        };
        for (size_t q = 0; q < idx.size(); q++)
            if (outputtable(idx[q])) (void)slot_of(hs[q], true);
        std::vector<std::vector<size_t>> lists(cap);
        for (size_t q = 0; q < idx.size(); q++) {
            const long j = slot_of(hs[q], false);
            if (j >= 0) lists[(size_t)j].push_back(q);
        }
        std::vector<uint8_t> kq((size_t)KL);
        for (const std::vector<size_t> &l : lists) {
            if (l.empty()) continue;
            std::vector<std::vector<uint8_t>> seen;   // distinct keys of this hash so far
            for (size_t q : l) {
                node_key(h, nodes[idx[q]].path, kq.data(), st.data());
                bool first = true;
                for (const auto &v : seen)
                    if (std::memcmp(v.data(), kq.data(), (size_t)KL) == 0) { first = false; break; }
                if (first) seen.push_back(kq);
                if (first && outputtable(idx[q])) keep[idx[q]] = 1;
            }
        }
    }
    long total = 0;
    for (size_t i = 0; i < nodes.size(); i++) {
        if (!keep[i] || nodes[i].dead) continue;
        const uint32_t solving = nodes[i].mask & nodes[i].live & 0xFEu;
        uint16_t insns[PCAGPU_MAX_KDEPTH];
        node_insns(h, nodes[i].path, insns, nullptr);
        for (int r = 1; r < 8; r++) {
            if (!(solving >> r & 1u)) continue;
            if (out && total < cap) {
                pca_solution_t &so = out[total];
                std::memset(&so, 0, sizeof(so));
                so.prefix_rank = h->rank_of(nodes[i].path.prefix);
                int n = 0;
                for (int q = 0; q < h->task.num_inputs; q++)
                    so.prog[n++] = enc_i(OP_IN, (uint32_t)q + 1, h->task.input_ports[q]);
                for (int q = 0; q < D; q++) so.prog[n++] = insns[q];
                so.prog[n++] = enc_i(OP_OUT, (uint32_t)r, h->task.output_ports[0]);
                so.prog[n++] = enc_r(OP_HLT, 0, 0, 0);
                so.len = (uint8_t)n;
                so.out_reg = (uint8_t)r;
                so.kernel_len = (uint8_t)D;
                uint32_t live = 0;
                for (int q = 0; q < h->task.num_inputs; q++) live |= 1u << (q + 1);
                for (int q = 0; q < D; q++) {
                    uint16_t cands[CAND_LMAX + 64];
                    int nc = gen_candidates(cands, CAND_LMAX + 64, live, q, D);
                    int idx = 0;
                    while (idx < nc && cands[idx] != insns[q]) idx++;
                    so.cand[q] = (uint16_t)idx;
                    live = live_after(live, insns[q]);
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
    const int rc = settle_scan(h);
    if (rc) return rc;
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

namespace {
uint64_t host_dg_mix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
}  // namespace

extern "C" int pcagpu_alu_digest(pcagpu_t *h, int op, uint32_t b_lo, uint32_t b_hi, uint64_t *out)
{
    if (!out || b_lo > b_hi || b_hi > 65536 || op < (int)OP_ADD || op > (int)OP_CMP) return PCAGPU_EINVAL;
    if (h) {
        const int rc = settle_scan(h);
        if (rc) return rc;
    }
    if (!h) {
        uint64_t acc = 0;
        for (uint32_t b = b_lo; b < b_hi; b++)
            for (uint32_t a = 0; a < 65536; a++) {
                uint32_t f = 0, v;
                bool fault = false;
                if ((uint32_t)op == OP_CMP) v = 0, f = flags_sub(a, b);
                else v = alu_eval((uint32_t)op, a, b, 0, 0, &f, &fault);
                if (fault) v = 0, f = 0;
                acc += host_dg_mix((uint64_t)a | ((uint64_t)b << 16) | ((uint64_t)v << 32) | ((uint64_t)f << 48) |
                                   ((uint64_t)fault << 56));
            }
        *out = acc;
        return 0;
    }
    (void)hipSetDevice(h->device);
    HIP_TRY(hipMemsetAsync(h->d_counters.p + 6, 0, sizeof(unsigned long long), h->stream));
    launch_alu_digest((uint32_t)op, b_lo, b_hi, h->d_counters.p + 6, h->stream);
    HIP_TRY(hipGetLastError());
    unsigned long long v = 0;
    HIP_TRY_S(h->stream, hipMemcpyAsync(&v, h->d_counters.p + 6, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY_S(h->stream, hipStreamSynchronize(h->stream));
    *out = v;
    return 0;
}

extern "C" int pcagpu_get_stats(pcagpu_t *h, pcagpu_stats_t *st)
{
    if (!h || !st) return PCAGPU_EINVAL;
    const int rc = settle_scan(h);   // a launch in flight counts once collected
    if (rc) return rc;
    *st = h->stats;
    return 0;
}

extern "C" void pcagpu_reset_stats(pcagpu_t *h)
{
    if (!h) return;
    (void)settle_scan(h);
    std::memset(&h->stats, 0, sizeof(h->stats));
}

// ------------------------------------------------------------ wavefront
namespace pca {

int expand_level(pcagpu *h, const DevTask &dt, const CandDev &cd, int cur, uint64_t n, int D, int pos,
                 uint64_t *kept_out, uint64_t max_children)
{
    int rc;
    ExpandArgs a;
    std::memset(&a, 0, sizeof(a));
    a.parents = h->d_front[cur].p;
    a.n_parents = n;
    a.D = D;
    a.pos = pos;
    a.cand = cd;
    a.key_mask = (h->opts.test_flags & 2u) ? 7ull : ~0ull;
    if ((rc = h->d_child_off.ensure(n + 1)) || (rc = h->d_keep_cnt.ensure(n + 1)) ||
        (rc = h->d_scan_tmp.ensure(exclusive_scan_tmp_words(n) + 16)))
        return rc;
    HIP_TRY(hipMemsetAsync(h->d_child_off.p, 0, (n + 1) * sizeof(uint64_t), h->stream));
    HIP_TRY(hipEventRecord(h->ev0, h->stream));
    launch_expand_count(a, h->d_child_off.p, h->stream);
    launch_exclusive_scan(h->d_child_off.p, n, h->d_scan_tmp.p, h->stream);
    uint64_t children = 0;
    HIP_TRY_S(h->stream, hipMemcpyAsync(&children, h->d_child_off.p + n, sizeof(uint64_t), hipMemcpyDeviceToHost,
                                        h->stream));
    HIP_TRY_S(h->stream, hipStreamSynchronize(h->stream));
    if (max_children && children > max_children) {
        *kept_out = 0;
        return 1;   // not expanded: d_front[cur] is untouched
    }
    uint64_t tcap = 1024;
    while (tcap < 2 * children) tcap <<= 1;
    if (h->opts.test_flags & 1u) tcap = 64;
    if ((rc = h->d_table.ensure(2 * tcap)) || (rc = h->d_keep_bits.ensure(children / 32 + 8)) ||
        (rc = h->d_child_hash.ensure(children ? children : 1)) || (rc = h->d_parent_rows.ensure(n * MAXT)))
        return rc;
    // keys 0 (empty), min slots all-ones so atomicMin keeps the smallest index
    launch_table_init(h->d_table.p, tcap, h->stream);
    HIP_TRY(hipMemsetAsync(h->d_keep_bits.p, 0, (children / 32 + 8) * sizeof(uint32_t), h->stream));
    HIP_TRY(hipMemsetAsync(h->d_counters.p + 4, 0, 2 * sizeof(unsigned long long), h->stream));
    a.child_off = h->d_child_off.p;
    a.table = h->d_table.p;
    a.table_mask = tcap - 1;
    a.child_hash = h->d_child_hash.p;
    a.parent_rows = h->d_parent_rows.p;
    a.keep_bits = h->d_keep_bits.p;
    a.keep_count = h->d_keep_cnt.p;
    a.counters = h->d_counters.p + 4;
    a.work = h->d_work.p;
    const int grid = grid_for_aux(h, n, AuxKernel::ExpandCompact);
    // insert / survive: slice the child lists of a small frontier so that
    // every resident wave has work, and take runs of units from a large one
    // (the runs sized by the insert kernel's grid, which survive shares: the
    // work counter's units must mean the same for both)
    const uint64_t resident = (uint64_t)grid_for_aux(h, UINT64_MAX / 2, AuxKernel::ExpandInsert) * SCAN_WAVES;
    a.slices = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(resident / std::max<uint64_t>(n, 1), 1), 16);
    const int grid_i = grid_for_aux(h, n * a.slices, AuxKernel::ExpandInsert);
    const int grid_v = grid_for_aux(h, n * a.slices, AuxKernel::ExpandSurvive);
    a.grab = (uint32_t)std::min<uint64_t>(
        std::max<uint64_t>(n * a.slices / ((uint64_t)grid_i * SCAN_WAVES * 8), 1), 32);
    HIP_TRY(hipMemsetAsync(h->d_work.p, 0, sizeof(unsigned long long), h->stream));
    launch_expand_insert(dt, a, grid_i, h->stream);
    HIP_TRY(hipMemsetAsync(h->d_work.p, 0, sizeof(unsigned long long), h->stream));
    HIP_TRY(hipMemsetAsync(h->d_keep_cnt.p, 0, (n + 1) * sizeof(uint64_t), h->stream));
    launch_expand_survive(dt, a, grid_v, h->stream);
    launch_exclusive_scan(h->d_keep_cnt.p, n, h->d_scan_tmp.p, h->stream);
    uint64_t kept = 0;
    unsigned long long ctr[2] = {0, 0};
    HIP_TRY_S(h->stream, hipMemcpyAsync(&kept, h->d_keep_cnt.p + n, sizeof(uint64_t), hipMemcpyDeviceToHost,
                                        h->stream));
    HIP_TRY_S(h->stream, hipMemcpyAsync(ctr, h->d_counters.p + 4, 2 * sizeof(unsigned long long),
                                        hipMemcpyDeviceToHost, h->stream));
    HIP_TRY_S(h->stream, hipStreamSynchronize(h->stream));
    if (ctr[0] || ctr[1]) {
        h->stats.oep_collisions += ctr[0];
        return ctr[0] ? PCAGPU_ECOLLISION : PCAGPU_EOVERFLOW;
    }
    const int nxt = cur ^ 1;
    if ((rc = h->d_front[nxt].ensure(kept ? kept : 1))) return rc;
    a.keep_off = h->d_keep_cnt.p;
    a.out = h->d_front[nxt].p;
    HIP_TRY(hipMemsetAsync(h->d_work.p, 0, sizeof(unsigned long long), h->stream));
    a.grab = (uint32_t)std::min<uint64_t>(std::max<uint64_t>(n / ((uint64_t)grid * SCAN_WAVES * 8), 1), 32);
    launch_expand_compact(a, grid, h->stream);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventRecord(h->ev1, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->stats.expand_ns += (uint64_t)((double)ms * 1e6);
    h->stats.frontier_nodes += children;
    h->stats.frontier_kept += kept;
    *kept_out = kept;
    return 0;
}

// The exact-OEP frontier of the depth-D tree, level by level from the root
// down to split s_max, into d_front[*cur]; stops early (*s_done < s_max) at
// the first level whose child lists would exceed max_children (0: no cap).
// Returns the record count or an error.
long gpu_frontier(pcagpu *h, int D, int s_max, uint64_t max_children, int *s_done, int *cur_out)
{
    int rc;
    pca_prefix_t root;
    std::memset(&root, 0, sizeof(root));
    for (int i = 0; i < h->task.num_inputs; i++) root.live |= (uint8_t)(1u << (i + 1));
    if ((rc = h->d_front[0].ensure(1))) return rc;
    HIP_TRY(hipMemcpy(h->d_front[0].p, &root, sizeof(root), hipMemcpyHostToDevice));
    uint64_t n = 1;
    int cur = 0, pos = 0;
    const CandDev cd = cand_dev(h);
    for (; pos < s_max && n > 0; pos++) {
        uint64_t kept = 0;
        if ((rc = expand_level(h, h->dt, cd, cur, n, D, pos, &kept, max_children)) < 0) return rc;
        if (rc == 1) break;
        n = kept;
        cur ^= 1;
    }
    *s_done = n > 0 ? pos : s_max;
    *cur_out = cur;
    return (long)n;
}

}  // namespace pca

extern "C" long pcagpu_frontier(pcagpu_t *h, int D, int s, pca_prefix_t *out, long cap)
{
    if (!h) return PCAGPU_EINVAL;
    if (D < 0 || D > PCAGPU_MAX_KDEPTH || s < 0 || s > D) return PCAGPU_EINVAL;
    (void)hipSetDevice(h->device);
    const int rc0 = settle_scan(h);
    if (rc0) return rc0;
    int s_done = 0, cur = 0;
    const long nl = gpu_frontier(h, D, s, 0, &s_done, &cur);
    if (nl < 0) return nl;
    const uint64_t n = (uint64_t)nl;
    if (out && n) {
        const uint64_t m = std::min<uint64_t>(n, (uint64_t)std::max<long>(cap, 0));
        if (m) HIP_TRY(hipMemcpy(out, h->d_front[cur].p, m * sizeof(pca_prefix_t), hipMemcpyDeviceToHost));
    }
    return (long)n;
}

int pca::hold_device_frontier(pcagpu *h, int cur, size_t n)
{
    if (!h->in_depth) return PCAGPU_ESTATE;
    if (n > h->front_cap) {
        if (h->h_front) (void)hipHostFree(h->h_front);
        h->h_front = nullptr;
        h->front_cap = 0;
        if (hipHostMalloc((void **)&h->h_front, n * sizeof(pca_prefix_t)) != hipSuccess) return PCAGPU_ENOMEM;
        h->front_cap = n;
    }
    if (n) HIP_TRY(hipMemcpy(h->h_front, h->d_front[cur].p, n * sizeof(pca_prefix_t), hipMemcpyDeviceToHost));
    h->dev_front = true;
    h->d_src = h->d_front[cur].p;
    h->n_front = n;
    return 0;
}

void pca::release_wavefront_scratch(pcagpu *h, int cur)
{
    h->d_table.release();
    h->d_child_hash.release();
    h->d_parent_rows.release();
    h->d_keep_bits.release();
    h->d_front[cur ^ 1].release();
}

// ------------------------------------------------------------ IDDFS
extern "C" int pcagpu_solve(pcagpu_t *h, const pcagpu_solve_opts_t *o, pcagpu_result_t *res,
                            pca_solution_t *out, long cap)
{
    if (!h || !o || !res) return PCAGPU_EINVAL;
    std::memset(res, 0, sizeof(*res));
    (void)hipSetDevice(h->device);
    const int rc0 = settle_scan(h);
    if (rc0) return rc0;
    res->optimal_depth = -1;
    res->depths_done = -1;
    const double t0 = now_s();
    h->deadline = o->time_limit > 0 ? t0 + o->time_limit : 0;
    h->stopped = false;
    struct ClearDeadline {
        pcagpu *h;
        ~ClearDeadline() { h->deadline = 0; }
    } clear_deadline{h};
    const int maxd = std::min(o->max_depth, PCAGPU_MAX_PROG - 2 - h->task.num_inputs);
    for (int D = std::max(o->start_depth, 0); D <= maxd && D <= PCAGPU_MAX_KDEPTH; D++) {
        int s = o->split < 0 ? (D > 0 ? D - 1 : 0) : std::min(o->split, D);
        int rc;
        const uint64_t cand0 = h->stats.candidates, pre0 = h->stats.prefixes;
        bool stopped = false;
        const double td = now_s();
        double t_front = td, t_copy = td;   // phase ends (verbose breakdown)
        if (o->producer == 0) {
            // GPU wavefront to the split; with the automatic split it stops at
            // the deepest level whose child lists stay within kAutoChildren
            // (the split only moves work between the frontier and the scan;
            // the solutions are the same)
            (void)hipSetDevice(h->device);
            int cur = 0, s_done = s;
            const uint64_t fn0 = h->stats.frontier_nodes, fk0 = h->stats.frontier_kept;
            const long n = gpu_frontier(h, D, s, o->split < 0 ? kAutoChildren : 0, &s_done, &cur);
            res->prefix_nodes[D] = h->stats.frontier_nodes - fn0;
            res->prefix_pruned[D] = res->prefix_nodes[D] - (h->stats.frontier_kept - fk0);
            if (n < 0) return (int)n;
            s = s_done;
            t_front = now_s();
            if ((rc = pcagpu_begin_depth(h, D, s))) return rc;
            if ((rc = hold_device_frontier(h, cur, (size_t)n))) return rc;
            t_copy = now_s();
            for (long i = 0; i < n; i += h->batch_cap) {
                const uint32_t c = (uint32_t)std::min<long>(h->batch_cap, n - i);
                h->batch_first = (uint64_t)i;
                h->stats.prefixes += c;
                if ((rc = scan_batch(h, h->h_front + i, h->d_src + i, c))) return rc;
                if (h->stopped || (o->time_limit > 0 && now_s() - t0 > o->time_limit)) { stopped = true; break; }
            }
        } else {
            if ((rc = pcagpu_begin_depth(h, D, s))) return rc;
            uint64_t rank = 0;
            DfsStats ds;
            int err = 0;
            dfs_prefixes(&h->task, D, s,
                         [&](const pca_prefix_t &p) {
                             int r = pcagpu_submit(h, &p, 1, rank++);
                             if (r) { err = r; return false; }
                             if (h->stopped) { stopped = true; return false; }
                             if (o->time_limit > 0 && (rank & 4095) == 0 && now_s() - t0 > o->time_limit) {
                                 stopped = true;
                                 return false;
                             }
                             return true;
                         },
                         &ds);
            if (err) return err;
            res->prefix_nodes[D] = ds.nodes;
            res->prefix_pruned[D] = ds.pruned;
        }
        const double t_sub = now_s();
        long nsol = pcagpu_end_depth(h, out, cap);
        if (nsol < 0) return (int)nsol;
        res->candidates[D] = h->stats.candidates - cand0;
        res->prefixes[D] = h->stats.prefixes - pre0;
        const double t_end = now_s();
        res->depth_seconds[D] = t_end - td;
        if (o->verbose) {
            std::printf("  depth %d (split %d): %llu prefixes, %llu candidates, %ld solutions, %.3f s\n", D, s,
                        (unsigned long long)res->prefixes[D], (unsigned long long)res->candidates[D], nsol,
                        t_end - t0);
            std::printf("    phases: frontier %.4f s, copy %.4f s, submit %.4f s, end of depth %.4f s\n",
                        t_front - td, t_copy - t_front, t_sub - t_copy, t_end - t_sub);
        }
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
