// noalloc.hip -- pcagpu_raw_hits_noalloc: the raw-hit scan of one prefix
// batch (what pcagpu_begin_depth + pcagpu_submit + pcagpu_flush +
// pcagpu_raw_hits compute) for callers that cannot allocate device memory.
//
// Every device buffer the scan needs -- the grammar tables, the task record,
// the prefixes and scan blocks, the hit records and the launch counters --
// lives in one file-scope device array, `scratch`, of a fixed size.  Each GPU
// holds its own copy; its address is looked up once per device (with that
// device current) on the first call there, which also uploads the grammar
// tables into it by copy.  After that a call only copies its inputs in,
// clears the counters with a fill kernel, launches the scan and copies the
// hits out: no hipMalloc, hipFree or other allocating call.  Calls on one
// GPU share its array, so calls that overlap on one GPU race; the host-side
// staging arrays below are shared the same way.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "pcagpu_impl.h"

namespace pca {
namespace {

constexpr uint32_t kNaPrefixes = PCAGPU_NOALLOC_MAX_PREFIXES;
constexpr uint32_t kNaBlocks = PCAGPU_NOALLOC_MAX_BLOCKS;
constexpr uint32_t kNaHits = PCAGPU_NOALLOC_MAX_HITS;

constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
// byte offsets of the parts of `scratch`
constexpr size_t kOffCtl = 0;   // u64: counters [0, 4), work counter [4], hit count (u32) [5]
constexpr int kCtlWords = 8;
constexpr size_t kOffTask = al256(kOffCtl + kCtlWords * sizeof(unsigned long long));
constexpr size_t kOffGrammar = al256(kOffTask + sizeof(DevTask));
constexpr size_t kOffPrefixes = al256(kOffGrammar + sizeof(GrammarDev));
constexpr size_t kOffBlocks = al256(kOffPrefixes + (size_t)kNaPrefixes * sizeof(pca_prefix_t));
constexpr size_t kOffHits = al256(kOffBlocks + (size_t)kNaBlocks * sizeof(ScanBlock));
constexpr size_t kScratchBytes = al256(kOffHits + (size_t)kNaHits * sizeof(HitRec));
static_assert(kScratchBytes < ((size_t)16 << 20), "scratch stays under 16 MiB");

}  // namespace
}  // namespace pca

// The device memory of every pcagpu_raw_hits_noalloc call on a device.
// Host code never uses this name as an address: it stands for the array
// only in hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[pca::kScratchBytes];

namespace pca {
namespace {

// the launch counters start at zero (a kernel, not a memset, so that a
// caller's stream capture would hold no memset node)
__global__ void na_clear_kernel(unsigned long long *p, int n)
{
    const int i = (int)threadIdx.x;
    if (i < n) p[i] = 0ull;
}

constexpr int kMaxDevices = 64;

struct NaDevice {
    unsigned char *ptr = nullptr;   // this device's `scratch`
    int cu = 0;
    int wpc[MAXLEV + 1] = {};       // resident scan waves per CU by state levels
    bool ready = false;
};

std::mutex g_mu;
NaDevice g_dev[kMaxDevices];
GrammarDev g_img;                   // host image of the grammar tables
bool g_img_ready = false;
// host staging of a call (shared like the device array)
ScanBlock g_blocks[kNaBlocks];
HitRec g_hits[kNaHits];
unsigned long long g_ctl[kCtlWords];

// Set-up on the first call on the current device: the array's address on
// this device, the grammar tables copied in, the device's CU count.
int na_device(NaDevice **out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return PCAGPU_ENODEV;
    if (dev < 0 || dev >= kMaxDevices) return PCAGPU_EINVAL;
    std::lock_guard<std::mutex> lock(g_mu);
    NaDevice &d = g_dev[dev];
    if (!d.ready) {
        unsigned char *ptr = nullptr;
        if (hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return PCAGPU_EHIP;
        if (!g_img_ready) {
            grammar_image(false, &g_img);
            g_img_ready = true;
        }
        HIP_TRY(hipMemcpy(ptr + kOffGrammar, &g_img, sizeof(GrammarDev), hipMemcpyHostToDevice));
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cu <= 0) cu = 256;
        d.ptr = ptr;
        d.cu = cu;
        d.ready = true;
    }
    *out = &d;
    return 0;
}

int na_grid(NaDevice *d, uint64_t units, uint32_t levels)
{
    const uint32_t lv = std::min<uint32_t>(std::max<uint32_t>(levels, 1u), (uint32_t)MAXLEV);
    if (!d->wpc[lv]) {
        int b = 0;
        d->wpc[lv] = (scan_occupancy(false, lv, &b) == hipSuccess && b > 0) ? b * SCAN_WAVES : 16;
    }
    const uint64_t waves = (uint64_t)d->cu * (uint64_t)d->wpc[lv];
    const uint64_t wgs = (std::min<uint64_t>(units, waves) + SCAN_WAVES - 1) / SCAN_WAVES;
    return (int)std::max<uint64_t>(wgs, 1);
}

}  // namespace
}  // namespace pca

using namespace pca;

extern "C" long pcagpu_raw_hits_noalloc(const pca_task_t *task, int target_depth, int split_depth,
                                        const pca_prefix_t *prefixes, uint32_t n, uint16_t *rows, long cap)
{
    if (!task || (n && !prefixes) || (cap > 0 && !rows)) return PCAGPU_EINVAL;
    if (task->num_outputs != 1 || task->num_inputs < 0 || task->num_inputs > 7 || task->num_tests < 0 ||
        task->num_tests > PCAGPU_MAX_TESTS || task->output_ports[0] > 7)
        return PCAGPU_EINVAL;
    for (int i = 0; i < task->num_inputs; i++)
        if (task->input_ports[i] > 7) return PCAGPU_EINVAL;
    const int D = target_depth, s = split_depth, k = D - s;
    if (D < 0 || D > PCAGPU_MAX_KDEPTH || s < 0 || s > D || task->num_inputs + D + 2 > PCAGPU_MAX_PROG)
        return PCAGPU_EINVAL;
    if (n > kNaPrefixes) return PCAGPU_EOVERFLOW;
    for (uint32_t i = 0; i < n; i++)
        if (prefixes[i].num_insns != s) return PCAGPU_EINVAL;
    NaDevice *d = nullptr;
    int rc = na_device(&d);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_mu);   // the host staging arrays
    // scan blocks of up to 64 children (k >= 2), or one unit per prefix
    const ScanTile tile{64, 2};
    uint32_t nb = 0;
    if (k >= 2)
        for (uint32_t i = 0; i < n; i++)
            if (!prefix_blocks(shared_grammar(), D, s, prefixes[i].live, i, g_blocks, kNaBlocks, &nb, tile.block_kids))
                return PCAGPU_EOVERFLOW;
    const uint64_t units = k >= 2 ? nb : n;
    DevTask dt;
    devtask_for(task, &dt);
    unsigned char *ptr = d->ptr;
    unsigned long long *ctl = reinterpret_cast<unsigned long long *>(ptr + kOffCtl);
    const hipStream_t st = nullptr;
    HIP_TRY_S(st, hipMemcpyAsync(ptr + kOffTask, &dt, sizeof(DevTask), hipMemcpyHostToDevice, st));
    if (n)
        HIP_TRY_S(st, hipMemcpyAsync(ptr + kOffPrefixes, prefixes, (size_t)n * sizeof(pca_prefix_t), hipMemcpyHostToDevice,
                               st));
    if (nb)
        HIP_TRY_S(st, hipMemcpyAsync(ptr + kOffBlocks, g_blocks, (size_t)nb * sizeof(ScanBlock), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(na_clear_kernel, dim3(1), dim3(64), 0, st, ctl, kCtlWords);
    HIP_TRY_S(st, hipGetLastError());
    uint32_t hits = 0;
    if (units) {
        ScanItem it;
        std::memset(&it, 0, sizeof(it));
        it.task = reinterpret_cast<const DevTask *>(ptr + kOffTask);
        it.prefixes = reinterpret_cast<const pca_prefix_t *>(ptr + kOffPrefixes);
        it.blocks = reinterpret_cast<const ScanBlock *>(ptr + kOffBlocks);
        it.n_prefix = n;
        it.n_units = units;
        it.D = D;
        it.s = s;
        it.k = k;
        it.tile = tile;
        it.hits = reinterpret_cast<HitRec *>(ptr + kOffHits);
        it.hit_count = reinterpret_cast<uint32_t *>(ctl + 5);
        it.hit_cap = kNaHits;
        ScanLaunch l;
        std::memset(&l, 0, sizeof(l));
        l.items = nullptr;
        l.n_items = 1;
        l.levels = scan_levels(k);
        l.n_units = units;
        l.cand = cand_dev(reinterpret_cast<const GrammarDev *>(ptr + kOffGrammar));
        l.counters = ctl;
        l.work = ctl + 4;
        l.tstat = nullptr;
        const int grid = na_grid(d, units, l.levels);
        l.grab = scan_grab(k, units, grid);
        launch_scan(l, &it, dt, grid, st);
        HIP_TRY_S(st, hipGetLastError());
        HIP_TRY_S(st, hipMemcpyAsync(g_ctl, ctl, sizeof(g_ctl), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        hits = (uint32_t)(g_ctl[5] & 0xFFFFFFFFull);
        if (hits > kNaHits) return PCAGPU_EOVERFLOW;
        if (hits) {
            HIP_TRY_S(st, hipMemcpyAsync(g_hits, ptr + kOffHits, (size_t)hits * sizeof(HitRec), hipMemcpyDeviceToHost, st));
            HIP_TRY(hipStreamSynchronize(st));
        }
    } else {
        HIP_TRY(hipStreamSynchronize(st));
    }
    // canonical order: prefix, then the suffix's candidate indices
    std::sort(g_hits, g_hits + hits, [k](const HitRec &a, const HitRec &b) {
        if (a.prefix != b.prefix) return a.prefix < b.prefix;
        for (int q = 0; q < k; q++)
            if (a.j[q] != b.j[q]) return a.j[q] < b.j[q];
        return false;
    });
    for (long i = 0; i < (long)hits && i < cap; i++) {
        uint16_t *r = rows + i * 14;
        std::memset(r, 0, 14 * sizeof(uint16_t));
        r[0] = (uint16_t)(g_hits[i].prefix & 0xFFFFu);
        r[1] = (uint16_t)(g_hits[i].prefix >> 16);
        r[4] = g_hits[i].flags;
        r[5] = g_hits[i].mask;
        for (int q = 0; q < k; q++) r[6 + q] = g_hits[i].j[q];
    }
    return (long)hits;
}
