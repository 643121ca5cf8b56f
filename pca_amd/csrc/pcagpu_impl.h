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

constexpr uint32_t kGroupMax = 1u << 20;   // scan groups per launch

}  // namespace pca

using pca::CandTable;
using pca::DevBuf;
using pca::DevTask;
using pca::HostHit;
using pca::HitRec;
using pca::MatchRec;
using pca::ScanGroup;
using pca::ScanItem;

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
    uint32_t waves_per_cu = 20;          // resident waves per CU, one-item scan launches
    uint32_t waves_per_cu_multi = 16;    // the same for multi-item launches

    DevBuf<uint16_t> d_lists;
    DevBuf<int> d_count, d_split;
    DevBuf<uint16_t> d_vg_insn, d_vg_j;
    DevBuf<uint8_t> d_vg_rdmask;
    DevBuf<int> d_vg_n;
    DevBuf<uint32_t> d_cand_total;
    DevBuf<pca_prefix_t> d_batch;
    DevBuf<ScanGroup> d_groups;
    DevBuf<uint64_t> d_goff;
    DevBuf<DevTask> d_task;          // the task record, resident for the kernels
    DevBuf<ScanItem> d_items;
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
    std::vector<int> order;          // device test order (index into task tests)

    // state-match pass buffers
    DevBuf<unsigned long long> d_q0, d_qf;
    DevBuf<MatchRec> d_match;
    DevBuf<uint32_t> d_match_count;
};


namespace pca {
int fill_devtask(pcagpu *h);
CandDev cand_dev(const pcagpu *h);
// persistent grid for `units` work units (multi: a launch of several items)
int grid_for(pcagpu *h, uint64_t units, bool multi = false);
size_t handle_device_bytes(const pcagpu *h);
// queue prefixes for the depth without flushing (the caller scans them)
int stage_prefixes(pcagpu *h, const pca_prefix_t *p, uint32_t n, uint64_t first_rank);
// group table of the staged batch into h_groups / h_goff; returns the group
// count (0 for k <= 1) and the unit count, or an error if it does not fit
int build_groups(pcagpu *h, uint32_t *n_groups, uint64_t *n_units);
}  // namespace pca
