// server.cpp -- continuous-batching serving of solve requests.
//
// A request is a task plus a depth budget; the answer is its proven optimal
// length and canonical program set (or "none up to the budget").  The
// server keeps up to max_slots tasks resident on the device (task record,
// prefix queue, hit buffer, wavefront/OEP buffers per slot, all counted
// against an HBM budget) and runs a scheduling loop:
//
//   admit queued requests into free slots while slots and budget allow
//   every active task contributes its next batch of split-depth prefixes,
//   its share of the step (prefixes_per_item, or more when few tasks are
//   active: step_prefixes split over them)
//   ONE persistent scan launch processes the batches of all tasks
//   tasks whose depth is complete are resolved on the host (exact OEP,
//   canonical order, CPU re-verification); solved ones free their slot,
//   the others move to the next depth
//
// Slots are opened ahead (warm_slots) and re-tasked per request, so a
// request pays no device allocation; a step zeroes and reads back one
// control block (counters, hit counts, test statistics) for all its
// launches, one host wait per step.
//
// Requests arriving while others are on the GPU join at the next step, so
// uneven tasks overlap in the same launches.  Each task's result is exactly
// what pcagpu_solve gives for it alone: the same prefixes, kernel and
// resolution, only scheduled together.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "host_dfs.h"
#include "host_pool.h"
#include "pcagpu_impl.h"

using namespace pca;

namespace {

// A finished depth with at most this many raw hits resolves in well under a
// millisecond (the heavy ones, millions of hits, take 0.1-0.2 s).
constexpr size_t kInlineRawHits = 1 << 14;

// Growable pinned host buffer (contents are not kept when it grows).
struct PinnedBuf {
    unsigned char *p = nullptr;
    size_t n = 0;
    int ensure(size_t want)
    {
        if (want <= n) return 0;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
        const size_t cap = std::max(want, (size_t)1 << 20);
        if (hipHostMalloc((void **)&p, cap) != hipSuccess) {
            p = nullptr;
            return PCAGPU_ENOMEM;
        }
        n = cap;
        return 0;
    }
    void release()
    {
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = 0;
    }
};

// What resolving a slot's finished depth led to.
struct Outcome {
    int rc = 0;          // < 0: the request fails
    bool done = false;   // optimum proven, or none up to the budget
    int optimal = -1;
    long nsol = 0;
    std::vector<pca_solution_t> sols;
    uint64_t ns = 0;     // host time the resolution took
};

struct Slot {
    pcagpu *h = nullptr;
    bool busy = false;
    uint64_t req = 0;
    int max_depth = 0;
    int D = 0;
    std::vector<pca_prefix_t> frontier;   // a host-produced depth's prefixes
    size_t next = 0;
    uint32_t staged = 0;     // prefixes of this step's launch
    size_t byte_cap = 0;     // device bytes this slot may keep between depths
    // a finished depth resolved in the background: 0 none, 1 running, 2 the
    // outcome is ready (res_state and res_out under pcagpu_server::res_mu);
    // while it runs, the step touches nothing of the slot but busy and h
    int res_state = 0;
    Outcome res_out;
    size_t bytes_seen = 0;   // device bytes when last counted outside a resolution
    // the depth's prefixes: the host-produced list, or the handle's
    // device-resident frontier (its pinned host view)
    bool dev() const { return h->dev_front; }
    size_t fsize() const { return h->dev_front ? h->n_front : frontier.size(); }
    const pca_prefix_t *fdata() const { return h->dev_front ? h->h_front : frontier.data(); }
};

struct Request {
    uint64_t id = 0;
    pca_task_t task;
    int max_depth = 0;
    int status = PCAGPU_REQ_QUEUED;
    int optimal_depth = -1;
    long nsol = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t prefixes = 0;
    std::vector<pca_solution_t> sols;
};

}  // namespace

struct pcagpu_server {
    pcagpu_server_opts_t opts;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    std::vector<Slot> slots;
    std::deque<uint64_t> queue;
    std::map<uint64_t, Request> reqs;
    std::vector<uint64_t> done_unpolled;
    uint64_t next_id = 1;
    uint32_t step_no = 0;   // steps that staged work (for opts.fail_step)
    double t0 = 0;
    DevBuf<ScanItem> d_items;
    // one step's control block, zeroed once and read back once: scan
    // counters [0,4), a work counter per launch, a 32-bit hit count per
    // item, then per launch the per-test counts of a one-item launch
    DevBuf<unsigned long long> d_ctl;
    unsigned long long *h_ctl = nullptr;   // pinned mirror
    // one step's prefixes and scan blocks of every item, copied in one go
    // (pinned host side), and the items' hit records, read back in one go
    PinnedBuf h_stage, h_hits;
    // the multi-item launch's item list, staged here (pinned, outliving the
    // launch) for its asynchronous upload
    PinnedBuf h_items;
    DevBuf<unsigned char> d_stage;
    std::unique_ptr<HostPool> pool;   // staging threads
    // finished depths are resolved on these threads while the steps go on
    // launching the other slots' work
    std::unique_ptr<JobQueue> resolver;
    std::mutex res_mu;
    std::condition_variable res_cv;
    std::vector<ScanItem> items;
    pcagpu_server_stats_t st;

    size_t max_launches() const { return (size_t)opts.max_slots + 1; }
    size_t hit_word() const { return 4 + max_launches(); }
    size_t tstat_word(size_t launch) const { return hit_word() + (opts.max_slots + 1) / 2 + launch * 2 * MAXT; }
    size_t ctl_words() const { return tstat_word(max_launches()); }
    unsigned long long *work_dev(size_t launch) { return d_ctl.p + 4 + launch; }
    uint32_t *hit_count_dev(size_t item) { return reinterpret_cast<uint32_t *>(d_ctl.p + hit_word()) + item; }
    uint32_t hit_count_host(size_t item) const { return reinterpret_cast<const uint32_t *>(h_ctl + hit_word())[item]; }
};

namespace {

double clock_s(const pcagpu_server *s) { return now_s() - s->t0; }

bool resolving(pcagpu_server *s, const Slot &sl)
{
    std::lock_guard<std::mutex> lk(s->res_mu);
    return sl.res_state != 0;
}

// (a slot being resolved counts with its bytes from before: its handle's
// buffers are the resolver's while it runs)
size_t slots_bytes(pcagpu_server *s)
{
    size_t b = 0;
    for (Slot &sl : s->slots) {
        if (!sl.h) continue;
        if (!resolving(s, sl)) sl.bytes_seen = handle_device_bytes(sl.h);
        b += sl.bytes_seen;
    }
    return b;
}

// The split-depth prefixes of depth D (split D-1, as pcagpu_solve's
// automatic split: the wavefront stops at the deepest level whose child
// lists fit kAutoChildren): a CPU DFS for splits of at most one level (at
// most a few hundred prefixes), the device wavefront beyond (the host
// thread also schedules every other task).
int start_depth(Slot &sl)
{
    pcagpu *h = sl.h;
    const int D = sl.D;
    int s = D > 0 ? D - 1 : 0;
    sl.frontier.clear();
    sl.next = 0;
    if (s <= 1) {
        long cap = 1 << 12, n;
        for (;;) {
            sl.frontier.resize((size_t)cap);
            n = pcagpu_host_frontier(&h->task, D, s, sl.frontier.data(), cap, nullptr, nullptr);
            if (n < 0) return (int)n;
            if (n < cap) break;
            cap *= 4;
        }
        sl.frontier.resize((size_t)n);
    } else {
        // the wavefront's last level stays on the device and is scanned in
        // place; the slot then gives back the expansion's large buffers if
        // it holds more than its share of the HBM budget
        int cur = 0, s_done = s;
        const long n = gpu_frontier(h, D, s, kAutoChildren, &s_done, &cur);
        if (n < 0) return (int)n;
        int rc = pcagpu_begin_depth(h, D, s_done);
        if (rc) return rc;
        if ((rc = hold_device_frontier(h, cur, (size_t)n))) return rc;
        if (sl.byte_cap && handle_device_bytes(h) > sl.byte_cap) release_wavefront_scratch(h, cur);
        return 0;
    }
    return pcagpu_begin_depth(h, D, s);
}

void finish(pcagpu_server *s, Slot &sl, int optimal, long nsol, const std::vector<pca_solution_t> &sols)
{
    Request &r = s->reqs[sl.req];
    r.status = PCAGPU_REQ_DONE;
    r.optimal_depth = optimal;
    r.nsol = nsol;
    r.sols = sols;
    r.t_done = clock_s(s);
    s->done_unpolled.push_back(r.id);
    s->st.completed++;
    sl.busy = false;
    sl.frontier.clear();
}

void fail_request(pcagpu_server *s, Slot &sl)
{
    Request &r = s->reqs[sl.req];
    r.status = PCAGPU_REQ_FAILED;
    r.t_done = clock_s(s);
    s->done_unpolled.push_back(r.id);
    s->st.failed++;
    sl.busy = false;
    sl.frontier.clear();
}

// Depth finished: resolve it and, while no program is found and budget is
// left, start the next depth.  Touches only the slot (its handle, stream and
// frontier), so finished slots are resolved on several host threads.
Outcome resolve_depths(Slot &sl)
{
    Outcome o;
    std::vector<pca_solution_t> buf(PCAGPU_SERVER_KEEP);
    for (;;) {
        const long n = pcagpu_end_depth(sl.h, buf.data(), (long)buf.size());
        if (n < 0) {
            o.rc = (int)n;
            return o;
        }
        if (n > 0) {
            buf.resize((size_t)std::min<long>(n, (long)buf.size()));
            o.done = true;
            o.optimal = sl.D;
            o.nsol = n;
            o.sols = std::move(buf);
            return o;
        }
        if (sl.D >= sl.max_depth) {
            o.done = true;
            return o;
        }
        sl.D++;
        const int rc = start_depth(sl);
        if (rc < 0) {
            o.rc = rc;
            return o;
        }
        if (sl.fsize()) return o;
        // an empty frontier closes the depth at once
    }
}

// The server-side effect of an outcome (single-threaded).
int apply(pcagpu_server *s, Slot &sl, const Outcome &o)
{
    if (o.rc < 0) return o.rc;
    if (o.done) finish(s, sl, o.optimal, o.nsol, o.sols);
    return 0;
}

int advance(pcagpu_server *s, Slot &sl) { return apply(s, sl, resolve_depths(sl)); }

// Resolves a slot's finished depth on a resolver thread.
void resolve_in_background(pcagpu_server *s, Slot &sl)
{
    {
        std::lock_guard<std::mutex> lk(s->res_mu);
        sl.res_state = 1;
    }
    Slot *p = &sl;   // slots never move: the vector holds max_slots from the start
    s->resolver->submit([s, p] {
        const double t = now_s();
        Outcome o = resolve_depths(*p);
        o.ns = (uint64_t)((now_s() - t) * 1e9);
        std::lock_guard<std::mutex> lk(s->res_mu);
        p->res_out = std::move(o);
        p->res_state = 2;
        s->res_cv.notify_all();
    });
}

// Applies the outcomes of the background resolutions that have finished
// (on the calling thread, the one that steps the server).
void apply_ready(pcagpu_server *s)
{
    for (Slot &sl : s->slots) {
        Outcome o;
        {
            std::lock_guard<std::mutex> lk(s->res_mu);
            if (sl.res_state != 2) continue;
            o = std::move(sl.res_out);
            sl.res_out = Outcome();
            sl.res_state = 0;
        }
        s->st.resolve_ns += o.ns;
        if (apply(s, sl, o) < 0) fail_request(s, sl);
    }
}

// f(0..n-1) on the server's host threads (the calling one included; only
// it when `parallel` is false).
template <class F>
void run_parallel(pcagpu_server *s, size_t n, F &&f, bool parallel = true)
{
    s->pool->run(n, std::forward<F>(f), parallel);
}

// A new resident slot on s's device, holding `task` (the caller re-tasks
// it per request).  Its prefix batch holds one step's largest share.
int open_slot(pcagpu_server *s, const pca_task_t *task, Slot **out)
{
    pcagpu_opts_t o;
    std::memset(&o, 0, sizeof(o));
    o.device = s->opts.device;
    o.batch_capacity = std::max(s->opts.prefixes_per_item, s->opts.step_prefixes);
    o.hit_capacity = s->opts.hit_capacity ? s->opts.hit_capacity : 1u << 16;
    pcagpu *h = nullptr;
    const int rc = pcagpu_open(task, &o, &h);
    if (rc < 0) return rc;
    s->slots.push_back(Slot());
    s->slots.back().h = h;
    s->slots.back().byte_cap = (size_t)(s->opts.hbm_budget / std::max<uint32_t>(s->opts.max_slots, 1u));
    s->st.slots_opened++;
    *out = &s->slots.back();
    return 0;
}

// True if one more slot fits max_slots and the HBM budget (the first slot
// always does, so a small budget serialises instead of stalling).
bool slot_fits(pcagpu_server *s)
{
    if (s->slots.size() >= s->opts.max_slots) return false;
    if (s->slots.empty()) return true;
    const size_t per = slots_bytes(s) / s->slots.size();
    return slots_bytes(s) + per <= s->opts.hbm_budget;
}

int admit(pcagpu_server *s)
{
    while (!s->queue.empty()) {
        Slot *free_slot = nullptr;
        for (Slot &sl : s->slots)
            if (!sl.busy && sl.h) { free_slot = &sl; break; }
        if (!free_slot) {
            if (!slot_fits(s)) return 0;
            Request &r = s->reqs[s->queue.front()];
            int rc = open_slot(s, &r.task, &free_slot);
            if (rc == PCAGPU_EINVAL) {
                // the task itself is unsupported: fail this request only
                const uint64_t bad = s->queue.front();
                s->queue.pop_front();
                Request &b = s->reqs[bad];
                b.status = PCAGPU_REQ_FAILED;
                b.t_done = clock_s(s);
                s->done_unpolled.push_back(bad);
                s->st.failed++;
                continue;
            }
            if (rc < 0) return rc;
        }
        const uint64_t id = s->queue.front();
        s->queue.pop_front();
        Request &r = s->reqs[id];
        Slot &sl = *free_slot;
        if (pcagpu_retask(sl.h, &r.task) < 0) {
            sl.req = id;
            fail_request(s, sl);
            continue;
        }
        sl.busy = true;
        sl.req = id;
        sl.max_depth = r.max_depth;
        sl.D = 0;
        r.status = PCAGPU_REQ_ACTIVE;
        r.t_admit = clock_s(s);
        s->st.admitted++;
        int rc = start_depth(sl);
        if (rc < 0) { fail_request(s, sl); continue; }
        if (!sl.fsize() && (rc = advance(s, sl)) < 0) fail_request(s, sl);
    }
    return 0;
}

}  // namespace

extern "C" int pcagpu_server_open(const pcagpu_server_opts_t *opts, pcagpu_server_t **out)
{
    if (!out) return PCAGPU_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return PCAGPU_ENODEV;
    pcagpu_server *s = new (std::nothrow) pcagpu_server();
    if (!s) return PCAGPU_ENOMEM;
    std::memset(&s->opts, 0, sizeof(s->opts));
    if (opts) s->opts = *opts;
    if (!s->opts.max_slots) s->opts.max_slots = 64;
    if (!s->opts.hbm_budget) s->opts.hbm_budget = (uint64_t)200 << 30;
    if (!s->opts.prefixes_per_item) s->opts.prefixes_per_item = 1u << 14;
    if (!s->opts.step_prefixes) s->opts.step_prefixes = 1u << 18;
    const uint32_t warm = s->opts.warm_slots == 0xFFFFFFFFu ? 0u
                          : std::min(s->opts.warm_slots ? s->opts.warm_slots : 4u, s->opts.max_slots);
    std::memset(&s->st, 0, sizeof(s->st));
    s->t0 = now_s();
    if (hipSetDevice(s->opts.device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&s->ev0) != hipSuccess || hipEventCreate(&s->ev1) != hipSuccess) {
        pcagpu_server_close(s);
        return PCAGPU_EHIP;
    }
    if (s->d_items.ensure(s->opts.max_slots) || s->d_ctl.ensure(s->ctl_words()) ||
        hipHostMalloc((void **)&s->h_ctl, s->ctl_words() * sizeof(unsigned long long)) != hipSuccess) {
        pcagpu_server_close(s);
        return PCAGPU_ENOMEM;
    }
    // up to 8 host threads in all (the caller is one of them)
    const int dev = s->opts.device;
    s->pool.reset(new (std::nothrow) HostPool(
        (unsigned)std::min<unsigned>(std::max(1u, std::thread::hardware_concurrency()), 8u) - 1u,
        [dev] { (void)hipSetDevice(dev); }));
    if (!s->pool) {
        pcagpu_server_close(s);
        return PCAGPU_ENOMEM;
    }
    s->resolver.reset(new (std::nothrow) JobQueue(std::min<uint32_t>(s->opts.max_slots, 4u),
                                                  [dev] { (void)hipSetDevice(dev); }));
    if (!s->resolver) {
        pcagpu_server_close(s);
        return PCAGPU_ENOMEM;
    }
    s->slots.reserve(s->opts.max_slots);
    // warm slots: opened now on a placeholder task (one input, one test) so
    // the first requests are only re-tasked
    pca_task_t ph;
    std::memset(&ph, 0, sizeof(ph));
    ph.num_inputs = 1;
    ph.num_outputs = 1;
    ph.input_ports[0] = 0;
    ph.output_ports[0] = 0;
    ph.num_tests = 1;
    for (uint32_t i = 0; i < warm && slot_fits(s); i++) {
        Slot *sl = nullptr;
        const int rc = open_slot(s, &ph, &sl);
        if (rc < 0) {
            pcagpu_server_close(s);
            return rc;
        }
    }
    *out = s;
    return 0;
}

extern "C" void pcagpu_server_close(pcagpu_server_t *s)
{
    if (!s) return;
    s->resolver.reset();   // waits for the resolutions in flight
    for (Slot &sl : s->slots) pcagpu_close(sl.h);
    s->d_items.release();
    s->d_ctl.release();
    if (s->h_ctl) (void)hipHostFree(s->h_ctl);
    s->h_stage.release();
    s->h_hits.release();
    s->h_items.release();
    s->d_stage.release();
    if (s->ev0) (void)hipEventDestroy(s->ev0);
    if (s->ev1) (void)hipEventDestroy(s->ev1);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

extern "C" double pcagpu_server_clock(pcagpu_server_t *s) { return s ? clock_s(s) : 0.0; }

extern "C" int pcagpu_server_submit(pcagpu_server_t *s, const pca_task_t *task, int max_depth, uint64_t *id)
{
    if (!s || !task || max_depth < 0 || max_depth > PCAGPU_MAX_KDEPTH) return PCAGPU_EINVAL;
    if (task->num_inputs + max_depth + 2 > PCAGPU_MAX_PROG) return PCAGPU_EINVAL;
    Request r;
    r.id = s->next_id++;
    r.task = *task;
    r.max_depth = max_depth;
    r.t_submit = clock_s(s);
    s->reqs[r.id] = r;
    s->queue.push_back(r.id);
    s->st.submitted++;
    if (id) *id = r.id;
    return 0;
}

extern "C" int pcagpu_server_step(pcagpu_server_t *s)
{
    if (!s) return PCAGPU_EINVAL;
    (void)hipSetDevice(s->opts.device);
    double tp = now_s();
    auto lap = [&](uint64_t &acc) {
        const double t = now_s();
        acc += (uint64_t)((t - tp) * 1e9);
        tp = t;
    };
    apply_ready(s);
    int rc = admit(s);
    lap(s->st.admit_ns);
    if (rc < 0) return rc;
    // ---- every active task stages its next batch: its share of the step
    // (at least prefixes_per_item, more when few tasks are active, so a task
    // running alone needs few steps)
    s->items.clear();
    uint64_t units = 0;
    uint32_t levels = 1;
    std::vector<Slot *> in_launch;
    uint32_t active = 0;
    // (a slot being resolved is neither staged nor looked at)
    std::vector<char> res_now(s->slots.size());
    for (size_t i = 0; i < s->slots.size(); i++) res_now[i] = resolving(s, s->slots[i]) ? 1 : 0;
    for (size_t i = 0; i < s->slots.size(); i++) {
        const Slot &sl = s->slots[i];
        active += !res_now[i] && sl.busy && sl.next < sl.fsize();
    }
    const uint32_t share = std::max(s->opts.prefixes_per_item, s->opts.step_prefixes / std::max(active, 1u));
    auto align16 = [](size_t b) { return (b + 15) & ~(size_t)15; };
    // staging is host work per slot (tile, prefix fitting, the batch and its
    // scan blocks): slots stage in parallel, items are assembled in slot order
    struct Staged {
        int rc = 0;
        uint32_t n = 0, nb = 0;
        uint64_t nu = 0;
        ScanTile tile{};
        const pca_prefix_t *dsrc = nullptr;   // prefixes already on the device
    };
    std::vector<Slot *> act;
    size_t left = 0;
    for (size_t i = 0; i < s->slots.size(); i++) {
        Slot &sl = s->slots[i];
        if (!res_now[i] && sl.busy && sl.next < sl.fsize()) {
            act.push_back(&sl);
            left += sl.fsize() - sl.next;
        }
    }
    std::vector<Staged> stg(act.size());
    // (threads only pay off with work to share: many slots or big batches)
    run_parallel(s, act.size(), [&](size_t i) {
        Slot &sl = *act[i];
        pcagpu *h = sl.h;
        Staged &g = stg[i];
        int cand;
        g.tile = current_tile(h, &cand);   // no tuning inside shared launches
        const pca_prefix_t *hp = sl.fdata() + sl.next;
        // a device-resident depth is read in place (nothing is staged but
        // its blocks): with single-prefix units it goes in one launch, with
        // blocks as many prefixes as one launch's blocks hold
        const size_t want = !sl.dev() ? std::min<size_t>(share, h->batch_cap)
                            : h->k <= 1 ? (size_t)0x7FFFFFFF
                                        : std::max<size_t>(share, h->batch_cap);
        uint32_t n = (uint32_t)std::min<size_t>(sl.fsize() - sl.next, want);
        n = std::max<uint32_t>(prefixes_fitting(h, hp, n, g.tile), 1);
        if (sl.dev()) {
            h->batch_first = sl.next;
            h->batch_n = n;
            h->stats.prefixes += n;
            g.dsrc = h->d_src + sl.next;
            if ((g.rc = build_blocks_of(h, hp, n, g.tile, &g.nb, &g.nu)) < 0) return;
        } else {
            if ((g.rc = stage_prefixes(h, hp, n, (uint64_t)sl.next)) < 0) return;
            if ((g.rc = build_blocks(h, g.tile, &g.nb, &g.nu)) < 0) return;
        }
        g.n = n;
    }, act.size() >= 8 || left >= (1u << 16));
    std::vector<uint32_t> staged_blocks;
    std::vector<const pca_prefix_t *> staged_src;
    size_t stage_bytes = 0;
    for (size_t i = 0; i < act.size(); i++) {
        Slot &sl = *act[i];
        pcagpu *h = sl.h;
        const Staged &g = stg[i];
        if (g.rc < 0) {
            // this task's batch could not be staged (out of memory, a bad
            // prefix): its request fails, the others go on; re-tasking the
            // slot resets its handle's batch
            fail_request(s, sl);
            continue;
        }
        staged_blocks.push_back(g.nb);
        staged_src.push_back(g.dsrc);
        stage_bytes += (g.dsrc ? 0 : align16((size_t)g.n * sizeof(pca_prefix_t))) +
                       align16((size_t)g.nb * sizeof(ScanBlock));
        levels = std::max(levels, scan_levels(h->k));
        ScanItem it;
        std::memset(&it, 0, sizeof(it));
        it.task = h->d_task.p;
        it.n_prefix = g.n;
        it.unit_begin = units;
        it.n_units = g.nu;
        it.D = h->D;
        it.s = h->s;
        it.k = h->k;
        it.tile = g.tile;
        it.hits = h->d_hits.p;
        it.hit_count = s->hit_count_dev(in_launch.size());
        it.hit_cap = h->hit_cap;
        s->items.push_back(it);
        units += g.nu;
        sl.staged = g.n;
        sl.next += g.n;
        in_launch.push_back(&sl);
    }
    // From here on the staged requests are in the launch: a failure fails
    // every one of them (their handles' batches are dropped with them, and
    // the stream is drained so that no copy of this step is still running
    // when the next step stages), and the server goes on with the rest.
    auto launch_and_collect = [&]() -> int {
    if (!s->items.empty()) {
        s->step_no++;
        if (s->opts.fail_step && s->step_no == s->opts.fail_step) return PCAGPU_ENOMEM;
    }
    // ---- every item's prefixes and blocks into one pinned buffer, one copy
    if (!s->items.empty()) {
        if ((rc = s->h_stage.ensure(stage_bytes)) || (rc = s->d_stage.ensure(s->h_stage.n))) return rc;
        size_t off = 0;
        for (size_t i = 0; i < s->items.size(); i++) {
            const pcagpu *h = in_launch[i]->h;
            const size_t bb = (size_t)staged_blocks[i] * sizeof(ScanBlock);
            if (staged_src[i]) {
                s->items[i].prefixes = staged_src[i];
            } else {
                const size_t pb = (size_t)s->items[i].n_prefix * sizeof(pca_prefix_t);
                std::memcpy(s->h_stage.p + off, h->h_batch, pb);
                s->items[i].prefixes = reinterpret_cast<const pca_prefix_t *>(s->d_stage.p + off);
                off += align16(pb);
            }
            if (bb) std::memcpy(s->h_stage.p + off, h->h_blocks, bb);
            s->items[i].blocks = reinterpret_cast<const ScanBlock *>(s->d_stage.p + off);
            off += align16(bb);
        }
        if (off) HIP_TRY(hipMemcpyAsync(s->d_stage.p, s->h_stage.p, off, hipMemcpyHostToDevice, s->stream));
    }
    // ---- the launches: an item whose units alone fill the chip several
    // times over runs in a launch of its own (the one-item kernel, its task
    // record in the kernel arguments); the others share one persistent
    // multi-item launch.  A unit of k >= 2 is a block of up to 64 parent
    // states, of k <= 1 a single one: sizes are compared in blocks.  The
    // launches go back to back on the stream; the host waits once, for the
    // control block's copy.
    if (!s->items.empty()) {
        const uint64_t resident = (uint64_t)grid_for(in_launch[0]->h, UINT64_MAX / 2, true, levels) * SCAN_WAVES;
        std::vector<ScanItem> shared;
        std::vector<size_t> alone, shared_idx;
        uint64_t shared_units = 0;
        uint32_t shared_levels = 1;
        int shared_k = 0;
        for (size_t i = 0; i < s->items.size(); i++) {
            const uint64_t blocks = s->items[i].k >= 2 ? s->items[i].n_units : s->items[i].n_units / 64;
            if (blocks >= 4 * resident) {
                alone.push_back(i);
            } else {
                ScanItem it = s->items[i];
                it.unit_begin = shared_units;
                shared_units += it.n_units;
                shared_levels = std::max(shared_levels, scan_levels(it.k));
                shared_k = std::max(shared_k, it.k);
                shared.push_back(it);
                shared_idx.push_back(i);
            }
        }
        if ((rc = s->d_items.ensure(std::max<size_t>(shared.size(), 1)))) return rc;
        lap(s->st.stage_ns);
        const size_t ctl_bytes = s->ctl_words() * sizeof(unsigned long long);
        HIP_TRY(hipMemsetAsync(s->d_ctl.p, 0, ctl_bytes, s->stream));
        HIP_TRY(hipEventRecord(s->ev0, s->stream));
        uint32_t n_launch = 0;
        // one-item launches count test outcomes per launch; the task's test
        // order adapts to them as in pcagpu_solve (multi-item launches mix
        // tasks and count nothing)
        std::vector<std::pair<uint32_t, Slot *>> counted;
        auto run = [&](const ScanItem *items, uint32_t n_items, uint64_t n_units, uint32_t lv, int k,
                       const DevTask &task0, Slot *one) -> int {
            if (n_items == 1) counted.emplace_back(n_launch, one);
            ScanLaunch l;
            l.items = s->d_items.p;
            l.n_items = n_items;
            l.levels = lv;
            l.n_units = n_units;
            l.cand = cand_dev(in_launch[0]->h);
            l.counters = s->d_ctl.p;
            l.tstat = n_items == 1 ? s->d_ctl.p + s->tstat_word(n_launch) : nullptr;
            l.work = s->work_dev(n_launch++);
            const int grid = grid_for(in_launch[0]->h, n_units, n_items > 1, lv);
            l.grab = scan_grab(k, n_units, grid);   // runs of units when they are single prefixes
            l.pad = 0;
            l.xcounters = nullptr;
            if (n_items > 1) {
                // never an asynchronous copy from pageable memory: the items
                // are a local vector, which an error return would free while
                // the copy is still queued
                const size_t ib = n_items * sizeof(ScanItem);
                if ((rc = s->h_items.ensure(ib))) return rc;
                std::memcpy(s->h_items.p, items, ib);
                HIP_TRY(hipMemcpyAsync(s->d_items.p, s->h_items.p, ib, hipMemcpyHostToDevice, s->stream));
            }
            launch_scan(l, items, task0, grid, s->stream);
            HIP_TRY(hipGetLastError());
            s->st.launches++;
            s->st.items += n_items;
            return 0;
        };
        for (size_t i : alone) {
            ScanItem it = s->items[i];
            it.unit_begin = 0;
            if ((rc = run(&it, 1, it.n_units, scan_levels(it.k), it.k, in_launch[i]->h->dt, in_launch[i])))
                return rc;
        }
        // (a single shared item also takes the one-item kernel, with its own
        // task record)
        if (!shared.empty() && (rc = run(shared.data(), (uint32_t)shared.size(), shared_units, shared_levels,
                                         shared_k, in_launch[shared_idx[0]]->h->dt, in_launch[shared_idx[0]])))
            return rc;
        HIP_TRY(hipEventRecord(s->ev1, s->stream));
        HIP_TRY(hipMemcpyAsync(s->h_ctl, s->d_ctl.p, s->tstat_word(n_launch) * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, s->stream));
        HIP_TRY(hipStreamSynchronize(s->stream));
        lap(s->st.wait_ns);
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, s->ev0, s->ev1));
        s->st.kernel_ns += (uint64_t)((double)ms * 1e6);
        s->st.candidates += s->h_ctl[0];
        s->st.suffixes += s->h_ctl[1];
        for (const auto &c : counted)
            if (!c.second->h->opts.fixed_test_order &&
                (rc = adapt_test_order(c.second->h, s->h_ctl + s->tstat_word(c.first))) < 0)
                return rc;
        // ---- collect each task's hits: every item's records copied into one
        // pinned buffer, one wait; an overflowing item rescans its batch
        // alone with a larger buffer
        size_t hit_total = 0;
        for (size_t i = 0; i < in_launch.size(); i++) {
            const uint32_t hits = s->hit_count_host(i);
            if (hits <= in_launch[i]->h->hit_cap) hit_total += hits;
        }
        if (hit_total) {
            if ((rc = s->h_hits.ensure(hit_total * sizeof(HitRec)))) return rc;
            size_t off = 0;
            for (size_t i = 0; i < in_launch.size(); i++) {
                const uint32_t hits = s->hit_count_host(i);
                if (!hits || hits > in_launch[i]->h->hit_cap) continue;
                HIP_TRY(hipMemcpyAsync(s->h_hits.p + off * sizeof(HitRec), in_launch[i]->h->d_hits.p,
                                       (size_t)hits * sizeof(HitRec), hipMemcpyDeviceToHost, s->stream));
                off += hits;
            }
            HIP_TRY(hipStreamSynchronize(s->stream));
        }
        size_t hit_off = 0;
        for (size_t i = 0; i < in_launch.size(); i++) {
            pcagpu *h = in_launch[i]->h;
            const uint32_t hits = s->hit_count_host(i);
            if (hits > h->hit_cap) {
                // overflow: this batch is rescanned alone (the flush grows the
                // hit buffer until it fits); if that fails, the request fails
                s->st.rescans++;
                if (pcagpu_flush(h) < 0) fail_request(s, *in_launch[i]);
                continue;
            }
            if (hits) {
                const HitRec *buf = reinterpret_cast<const HitRec *>(s->h_hits.p) + hit_off;
                hit_off += hits;
                for (uint32_t q = 0; q < hits; q++) {
                    const HitRec &r = buf[q];
                    HostHit x;
                    x.prefix = h->batch_first + r.prefix;
                    x.mask = r.mask;
                    x.flags = r.flags;
                    std::memcpy(x.j, r.j, sizeof(x.j));
                    h->raw.push_back(x);
                }
            }
            h->batch_first += h->batch_n;
            h->batch_n = 0;
        }
    }
    return 0;
    };
    if ((rc = launch_and_collect()) < 0) {
        (void)hipStreamSynchronize(s->stream);
        for (Slot *sl : in_launch)
            if (sl->busy) {
                sl->h->batch_n = 0;
                fail_request(s, *sl);
            }
        return rc;
    }
    lap(s->st.stage_ns);   // hit collection
    // ---- tasks whose depth is complete: resolved on the resolver threads
    // (exact OEP, CPU re-verification, the next depth's frontier) while the
    // next steps launch the other tasks' work; their outcomes are applied at
    // the start or end of a later step, in slot order.  Light depths (few
    // raw hits) are resolved here, on the step's host threads, and join the
    // next step: handed to a resolver they would miss it and wait a step
    // longer, for work that takes far less than a step.
    std::vector<Slot *> light;
    for (size_t i = 0; i < s->slots.size(); i++) {
        Slot &sl = s->slots[i];
        if (res_now[i] || !sl.busy || sl.next < sl.fsize()) continue;
        if (sl.h->raw.size() <= kInlineRawHits)
            light.push_back(&sl);
        else
            resolve_in_background(s, sl);
    }
    if (!light.empty()) {
        const double t = now_s();
        std::vector<Outcome> outs(light.size());
        run_parallel(s, light.size(), [&](size_t i) { outs[i] = resolve_depths(*light[i]); }, light.size() > 1);
        s->st.resolve_ns += (uint64_t)((now_s() - t) * 1e9);
        for (size_t i = 0; i < light.size(); i++)
            if (apply(s, *light[i], outs[i]) < 0) fail_request(s, *light[i]);
    }
    // with nothing launched this step, wait a little for a resolution
    // instead of returning at once to a caller that steps again
    if (s->items.empty()) {
        std::unique_lock<std::mutex> lk(s->res_mu);
        bool running = false;
        for (const Slot &sl : s->slots) running = running || sl.res_state == 1;
        if (running) s->res_cv.wait_for(lk, std::chrono::milliseconds(1));
    }
    apply_ready(s);
    lap(s->st.wait_ns);
    int pending = (int)s->queue.size();
    for (const Slot &sl : s->slots) pending += sl.busy;
    return pending;
}

extern "C" int pcagpu_server_poll(pcagpu_server_t *s, pcagpu_request_t *out, int cap)
{
    if (!s || (!out && cap > 0)) return PCAGPU_EINVAL;
    int n = 0;
    while (n < cap && !s->done_unpolled.empty()) {
        const Request &r = s->reqs[s->done_unpolled.front()];
        s->done_unpolled.erase(s->done_unpolled.begin());
        pcagpu_request_t &o = out[n++];
        std::memset(&o, 0, sizeof(o));
        o.id = r.id;
        o.status = r.status;
        o.optimal_depth = r.optimal_depth;
        o.num_solutions = r.nsol;
        o.max_depth = r.max_depth;
        o.t_submit = r.t_submit;
        o.t_admit = r.t_admit;
        o.t_done = r.t_done;
        if (!r.sols.empty()) o.first = r.sols[0];
    }
    return n;
}

extern "C" long pcagpu_server_solutions(pcagpu_server_t *s, uint64_t id, pca_solution_t *out, long cap)
{
    if (!s) return PCAGPU_EINVAL;
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return PCAGPU_EINVAL;
    const Request &r = it->second;
    if (r.status != PCAGPU_REQ_DONE) return PCAGPU_ESTATE;
    for (long i = 0; i < cap && i < (long)r.sols.size(); i++) out[i] = r.sols[(size_t)i];
    return r.nsol;
}

extern "C" int pcagpu_server_get_stats(pcagpu_server_t *s, pcagpu_server_stats_t *st)
{
    if (!s || !st) return PCAGPU_EINVAL;
    *st = s->st;
    size_t b = slots_bytes(s);
    st->device_bytes = b;
    st->slots = s->slots.size();
    return 0;
}
