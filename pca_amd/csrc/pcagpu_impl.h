// pcagpu_impl.h -- internal: the device handle behind pcagpu_t and the
// helpers the serving layer shares with the single-task driver.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <string>
#include <vector>

#include "host_core.h"
#include "kernels.h"

#define HIP_TRY(expr)                                  \
    do {                                               \
        if ((expr) != hipSuccess) return PCAGPU_EHIP;  \
    } while (0)

// For the calls of a sequence that reads back into host memory of the
// caller's frame (or a pageable buffer): on failure the stream is drained
// before returning, so that no copy already queued can still write into
// (or read from) memory that the return frees.
#define HIP_TRY_S(st, expr)                            \
    do {                                               \
        if ((expr) != hipSuccess) {                    \
            (void)hipStreamSynchronize(st);            \
            return PCAGPU_EHIP;                        \
        }                                              \
    } while (0)

namespace pca {

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
    size_t bytes() const { return n * sizeof(T); }
};

inline double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr uint32_t kGroupMax = 1u << 20;   // match-pass groups per launch
constexpr uint32_t kBlockMax = 1u << 21;   // scan blocks per launch

}  // namespace pca

using pca::CandDev;
using pca::CandTable;
using pca::GrammarDev;
using pca::DevBuf;
using pca::DevTask;
using pca::HostHit;
using pca::HitRec;
using pca::MatchRec;
using pca::ScanBlock;
using pca::ScanTile;
using pca::TILE_PER_CHILD;
using pca::TILE_SHARED;
using pca::ScanGroup;
using pca::ScanItem;

struct pcagpu {
    pca_task_t task;
    int T = 0;
    DevTask dt;
    const CandTable &ct = pca::shared_grammar();   // immutable, shared by all handles
    int device = 0;
    int cu_count = 256;
    std::string devname;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_up = nullptr;   // the pending scan launch's uploads are done
    pcagpu_opts_t opts;
    // The scan launch in flight (queued by a flush, not yet collected): its
    // batch stays in d_batch / d_blocks until it is collected, so an
    // overflowing launch can be redone; up_waited: h_batch and h_blocks may
    // be rewritten (the launch's uploads have finished).
    struct {
        bool on = false;
        const pca_prefix_t *src = nullptr;   // the launch's prefixes (device)
        uint32_t n = 0, n_blocks = 0;
        uint64_t units = 0;
        ScanTile tile{};
        int cand = -1;
        int k = 0;
        uint64_t batch_first = 0;
    } pend;
    bool up_waited = true;
    // pcagpu_solve's time limit (absolute, now_s(); 0: none): a flush stops
    // queueing launches once it has passed and sets `stopped`
    double deadline = 0;
    bool stopped = false;
    uint32_t waves_per_cu = 0;           // opts.waves_per_cu, or 0: from the kernel's occupancy
    // resident scan waves per CU by [multi-item][state levels], 0 = not yet asked
    uint32_t scan_wpc[2][pca::MAXLEV + 1] = {};
    uint32_t aux_wpc[(int)pca::AuxKernel::Count] = {};   // resident waves per CU of the frontier / match kernels

    const GrammarDev *d_grammar = nullptr;   // per device, shared, never freed
    const GrammarDev *d_grammar_bl = nullptr;   // branchless tables (leaf completion), same lifetime
    uint64_t complete_explored = 0;          // children generated by the last pcagpu_complete
    DevBuf<pca_prefix_t> d_batch;
    DevBuf<ScanBlock> d_blocks;
    DevBuf<ScanGroup> d_groups;      // match pass only, allocated on first use
    DevBuf<uint64_t> d_goff;
    DevBuf<DevTask> d_task;          // the task record, resident for the kernels
    DevBuf<ScanItem> d_items;
    DevBuf<HitRec> d_hits;
    DevBuf<uint32_t> d_hit_count;
    DevBuf<unsigned long long> d_counters, d_work;
    DevBuf<unsigned long long> d_tstat;   // [2 * MAXT] survivor-phase reach / end counts of a launch
    pca::TestState roots[PCAGPU_MAX_TESTS];   // per test (task order) state after the IN scaffold
    uint64_t t_reach[PCAGPU_MAX_TESTS] = {}, t_end[PCAGPU_MAX_TESTS] = {};   // by task test index
    // scan tile autotuner, per suffix length k: candidates tried so far,
    // their kernel ns per parent state, the winner (-1: still trying)
    int tile_tried[pca::MAXLEV + 1] = {};
    double tile_cost[pca::MAXLEV + 1][8] = {};
    int tile_best[pca::MAXLEV + 1] = {-1, -1, -1, -1, -1, -1, -1, -1, -1, -1};
    // graph-captured scan launch (see graph_scan); built on first use
    hipGraph_t g_graph = nullptr;
    hipGraphExec_t g_exec = nullptr;
    hipGraphNode_t g_prefixes = nullptr, g_blocks = nullptr, g_kernel = nullptr;
    void *h_out = nullptr;                  // pinned read-back area of a launch
    pca_prefix_t *h_batch = nullptr;
    ScanBlock *h_blocks = nullptr;
    uint32_t block_cap = 0;          // h_blocks / d_blocks records (kBlockMax once a k >= 2 scan ran)
    ScanGroup *h_groups = nullptr;
    uint64_t *h_goff = nullptr;
    uint32_t batch_cap = 0;
    uint32_t hit_cap = 0;

    // wavefront buffers
    DevBuf<pca_prefix_t> d_front[2];
    DevBuf<uint64_t> d_child_off, d_keep_cnt, d_scan_tmp;
    DevBuf<unsigned long long> d_table;
    DevBuf<uint32_t> d_keep_bits;
    DevBuf<uint64_t> d_child_hash;
    DevBuf<pca::TestState> d_parent_rows;

    // depth state
    bool in_depth = false;
    int D = 0, s = 0, k = 0;
    uint32_t batch_n = 0;
    uint64_t batch_first = 0;   // depth-prefix index of h_batch[0]
    std::vector<pca_prefix_t> prefixes;
    std::vector<uint64_t> ranks;
    // A device-resident depth (the GPU producer): the scan reads the
    // prefixes in place at d_src, h_front holds their host copy (pinned),
    // and prefix i has rank i.  Otherwise the prefixes are the submitted
    // ones in `prefixes` / `ranks`.
    bool dev_front = false;
    const pca_prefix_t *d_src = nullptr;
    pca_prefix_t *h_front = nullptr;
    size_t front_cap = 0, n_front = 0;
    const pca_prefix_t *pfx() const { return dev_front ? h_front : prefixes.data(); }
    size_t n_pfx() const { return dev_front ? n_front : prefixes.size(); }
    uint64_t rank_of(size_t i) const { return dev_front ? (uint64_t)i : ranks[i]; }
    std::vector<HostHit> raw;
    pcagpu_stats_t stats;
    std::vector<int> order;          // device test order (index into task tests)

    // state-match pass buffers
    DevBuf<unsigned long long> d_q0, d_qf;
    DevBuf<MatchRec> d_match;
    DevBuf<uint32_t> d_match_count;
};


namespace pca {
// Collects the scan launch in flight, if any: waits for it, accounts its
// counters and hits (redoing it with a larger hit buffer on overflow), and
// feeds the tile tuner and the test order.  Every entry point that reads
// results or uses the handle's stream for other work calls it first.
int settle_scan(pcagpu *h);
int fill_devtask(pcagpu *h);
// the device task record of a task in its initial device test order
void devtask_for(const pca_task_t *task, DevTask *dt);
// the grammar tables as uploaded to the device (host image)
void grammar_image(bool branchless, GrammarDev *img);
// Appends the scan blocks (k = D - s >= 2; bk children each) of prefix i,
// whose live mask is `live`, to out[*nb, cap); false, with *nb unchanged,
// if they do not fit.
bool prefix_blocks(const CandTable &ct, int D, int s, uint32_t live, uint32_t i, ScanBlock *out, uint32_t cap,
                   uint32_t *nb, int bk);
// device test order: h->dt's per-test fields and h->order from `order`
void set_test_order(pcagpu *h, const std::vector<int> &order);
// Folds one launch's per-position test counts (2 per position) into the
// task's statistics and re-sorts its test order; uploads the task record
// if the order changed.
int adapt_test_order(pcagpu *h, const unsigned long long *ts);
// the grammar tables of the current device, uploaded on first use and kept
// until the process ends (read-only, shared by every handle on the device)
const GrammarDev *shared_grammar_dev(int device, bool branchless = false);
CandDev cand_dev(const GrammarDev *g);
CandDev cand_dev(const pcagpu *h);
// One wavefront level: the children of the n records in d_front[cur] at
// kernel position pos (of depth D) on task record dt, first visitor of
// every distinct state kept (exact keys), written in rank order to
// d_front[cur ^ 1]; *kept = their count.
// With max_children > 0, a level whose child lists would be longer is not
// expanded: returns 1 and leaves d_front[cur] as it was.
int expand_level(pcagpu *h, const DevTask &dt, const CandDev &cd, int cur, uint64_t n, int D, int pos,
                 uint64_t *kept, uint64_t max_children = 0);
long gpu_frontier(pcagpu *h, int D, int s_max, uint64_t max_children, int *s_done, int *cur);
// child-list cap of one wavefront level under the automatic split (2^27:
// an OEP table of at most 2^29 slots, 8 GiB)
constexpr uint64_t kAutoChildren = 1ull << 27;
// persistent grid for `units` work units of a scan launch (multi: several
// items; levels: state levels per wave), or of the other persistent kernels
// (levels 0: the one-item scan figure at 3 levels)
int grid_for(pcagpu *h, uint64_t units, bool multi = false, uint32_t levels = 0);
// persistent grid of a frontier or match kernel, from that kernel's own occupancy
int grid_for_aux(pcagpu *h, uint64_t units, pca::AuxKernel which);
size_t handle_device_bytes(const pcagpu *h);
// queue prefixes for the depth without flushing (the caller scans them)
int stage_prefixes(pcagpu *h, const pca_prefix_t *p, uint32_t n, uint64_t first_rank);
// scan blocks of the staged batch into h_blocks (k >= 2); returns the block
// count and the unit count, or PCAGPU_EOVERFLOW if they do not fit one launch
int build_blocks(pcagpu *h, ScanTile tile, uint32_t *n_blocks, uint64_t *n_units);
// the same for n prefixes whose host view is hostp (a range of a
// device-resident depth)
int build_blocks_of(pcagpu *h, const pca_prefix_t *hostp, uint32_t n, ScanTile tile, uint32_t *n_blocks,
                    uint64_t *n_units);
// Makes the open depth device-resident: its n prefixes are d_front[cur]
// (read in place by the scans), copied once to the pinned host view h_front.
int hold_device_frontier(pcagpu *h, int cur, size_t n);
// Frees the wavefront's large buffers except the frontier d_front[cur]
// (table, keys, parent rows, survivor bits, the other level).
void release_wavefront_scratch(pcagpu *h, int cur);
// how many of the n prefixes p (not yet staged) have scan blocks that fit one
// launch's block buffer together (n when k <= 1)
uint32_t prefixes_fitting(const pcagpu *h, const pca_prefix_t *p, uint32_t n, ScanTile tile);
// the scan tile for the handle's current suffix length: fixed by the
// options, the tuner's winner, or (*cand >= 0) the candidate to try next
ScanTile current_tile(const pcagpu *h, int *cand);
// state levels per wave a scan at k suffix positions keeps in LDS
inline uint32_t scan_levels(int k) { return k < 1 ? 1u : (uint32_t)k; }
// units a wave takes per atomic on the scan's work counter: runs of about an
// eighth of a wave's share when units are single prefixes (k <= 1, cheap),
// one at a time when they are blocks of up to 64 parents each
inline uint32_t scan_grab(int k, uint64_t units, int grid)
{
    if (k >= 2) return 1;
    const uint64_t waves = (uint64_t)grid * SCAN_WAVES;
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(units / (waves * 8), 1), 32);
}
}  // namespace pca
