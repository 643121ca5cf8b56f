"""Benchmark of the MI355X search path (one JSON line on stdout).

Workload (the hybrid CPU-prefix / GPU-suffix configuration on 1 GPU):
is_power2 with 32 tests (16 powers of two, 0, 0xFFFF, 14 seeded non-powers),
IDDFS window at kernel depth 8 with split depth 5: the host DFS (exact OEP)
emits depth-5 prefixes in DFS order and the device evaluates every 3-long
suffix below them with every output register.

A step = one flush of a batch of B split-depth prefixes through the scan
kernel (inputs already resident in HBM when the timed region starts).
value = candidates/s, a candidate being one (prefix, suffix, out_reg) that
passes the grammar / liveness / dead-instruction filters (the same counter
as the CPU baseline's leaf tries).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B]
"""
import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_INT_OPS = 256 * 4 * 16 * 2.4e9     # gfx950 VALU lane-ops/s (256 CU x 4 SIMD x 16 lanes x 2.4 GHz)
OPS_PER_TEST = 9                         # one last-position VM step (8) + its target compare (1)
OPS_PER_PARENT_TEST = 15                 # parent step (8) + 7 out-register match compares


def cpu_baseline(task_path, seconds):
    """The oracle's restatement of the reference CPU solver (branch mode,
    32-bit OEP table as in tools/enumerate.c, single thread), timed over a
    bounded window of the same kernel-depth-8 search."""
    exe = os.path.join(ROOT, "oracle", "oracle_enumerate")
    if not os.path.exists(exe):
        return None
    out = subprocess.run([exe, task_path, "--start-depth", "8", "-d", "8", "--ref32", "--no-shadow",
                          "--time-limit", str(seconds), "--json"], capture_output=True, text=True)
    try:
        rec = json.loads(out.stdout.strip().splitlines()[-1])
    except Exception:
        return None
    secs = max(rec["seconds"], 1e-9)
    leaf = rec["leaf_tried"][8]
    tried = rec["tried"][8]
    return {"value": leaf / secs, "unit": "candidates/s", "cores": 1, "kind": "port",
            "sample": f"is_power2_32, kernel depth 8, first {secs:.1f} s of the branch-mode DFS "
                      f"(restated tools/enumerate.c, 1 thread of {os.cpu_count()}); "
                      f"reference tried counter {tried / secs:.4g}/s",
            "tried_per_s": tried / secs}


def serving_bench(rates=(10, 50, 200), n_req=200, seed=7):
    """Poisson arrivals of generated requests (depth budgets 4-8) at each
    rate, through the continuous-batching server and through the plain
    one-task-at-a-time GPU path on the same arrival schedule."""
    from pca_amd import taskgen, serving
    out = {"requests_per_rate": n_req, "task": "shallow_task (1-3 insn source, 1-2 inputs, 8-32 tests)",
           "budgets": "uniform 4-8"}
    for rate in rates:
        tasks = [taskgen.shallow_task(seed + rate, i) for i in range(n_req)]
        budgets = [4 + (i * 7 + rate) % 5 for i in range(n_req)]
        arr = serving.poisson_schedule(n_req, rate, seed=rate)
        recs, st = serving.poisson_run(tasks, budgets, arr, max_slots=64)
        print(f"serving rate {rate}: server done", file=sys.stderr, flush=True)
        seq = serving.sequential_run(tasks, budgets, arr)
        same = all(a[2:] == b[2:] for a, b in zip(recs, seq))
        span = max(r[1] for r in recs) - min(r[0] for r in recs)
        out[f"{rate}_req_s"] = {
            "server": serving.latency_stats(recs), "one_at_a_time": serving.latency_stats(seq),
            "server_candidates_per_s": st["candidates"] / max(span, 1e-9),
            "launches": st["launches"], "items_per_launch": st["items"] / max(st["launches"], 1),
            "same_results": same}
        print(f"serving rate {rate}: one-at-a-time done", file=sys.stderr, flush=True)
    return out


def suite_bench(n_tasks, budget, seed=11):
    """Config-3 batched suite: seeded random straight-line programs
    (taskgen.random_task), every task submitted at once to the server (one
    persistent launch per step over all active tasks) and, for comparison,
    one at a time through Solver.solve."""
    from pca_amd import taskgen, serving
    tasks = [taskgen.random_task(seed, i) for i in range(n_tasks)]
    # the slots are opened with the server (warm), as a long-lived server
    # would at start-up; the open time is reported beside the wall time
    a, wa, st = serving.suite_run(tasks, budget, max_slots=256, warm_slots=256)
    print("suite: server done", file=sys.stderr, flush=True)
    b, wb = serving.sequential_suite(tasks, budget)
    print("suite: one-at-a-time done", file=sys.stderr, flush=True)
    return {"tasks": n_tasks, "depth_budget": budget, "generator": f"taskgen.random_task(seed {seed})",
            "solved_within_budget": sum(1 for x in a if x[0] >= 0),
            "server_wall_s": wa, "server_open_s": st["open_s"], "one_at_a_time_wall_s": wb,
            "server_kernel_s": st["kernel_ns"] * 1e-9,
            "server_host_s": {k: st[k + "_ns"] * 1e-9 for k in ("admit", "stage", "wait", "resolve")},
            "launches": st["launches"], "items_per_launch": st["items"] / max(st["launches"], 1),
            "candidates": st["candidates"], "same_results": a == b}


def value_bench(batch=512, steps=30, seed=5):
    """Value network on the GPU (BF16 MFMA kernels): training and inference
    throughput in search states per second on synthetic states (random
    register files of is_power2_32's tests), and the layer-1 forward GEMM
    (M = 32 tests x batch, N = 512, K = 180) against the dense BF16 MFMA
    peak and against HBM, which bounds it at this K."""
    import numpy as np
    from pca_amd import taskgen
    from pca_amd._lib import lib
    from pca_amd.value import ValueNet, state_records, PER_TEST, H1, MAX_T
    rng = np.random.default_rng(seed)
    task = taskgen.is_power2_32(42)
    rec, _ = state_records(task, [[]] * batch)
    rec[:, :task.num_tests, 3:11] = rng.integers(0, 1 << 16, size=(batch, task.num_tests, 8))
    nvalid = np.full(batch, task.num_tests, dtype=np.uint8)
    depth = rng.integers(0, 8, size=batch).astype(np.int32)
    label = (rng.random(batch) < 0.3).astype(np.float32)
    net = ValueNet(seed=seed)
    X, nv, dp = net.upload(rec, nvalid, depth)

    def timed(fn, n):
        fn()
        lib().pcav_synchronize()
        t0 = time.perf_counter()
        for _ in range(n):
            fn()
        lib().pcav_synchronize()
        return (time.perf_counter() - t0) / n

    def train_step():
        net.forward(X, nv, dp)
        net.backward(label)
        net.adam_step()

    t_train = timed(train_step, steps)
    t_infer = timed(lambda: net.forward(X, nv, dp), steps)
    # the layer-1 GEMM at a large batch (4096 states: M = 131072 rows)
    big = 4096
    rec_b = np.repeat(rec, big // batch, axis=0)
    Xb, _, _ = net.upload(rec_b, np.repeat(nvalid, big // batch), np.repeat(depth, big // batch))
    h1 = net._buf("h1", (big * MAX_T, H1), np.uint16)
    t_gemm = timed(lambda: net._gemm(big * MAX_T, H1, PER_TEST, Xb, PER_TEST, 0, net.v16["W1"], H1, 0,
                                     bias=net.views["b1"], relu=True, C16=h1, ldc16=H1), 50)
    M = big * MAX_T
    flops = 2.0 * M * H1 * PER_TEST
    # algorithmic HBM bytes: A (M x K BF16) and W1 read once, C (M x N BF16)
    # written once; at K = 180 that is ~133 flop per byte, under the 312
    # flop/byte balance of 2.5 PFLOP/s over 8 TB/s, so HBM bounds it
    nbytes = 2.0 * (M * PER_TEST + PER_TEST * H1 + M * H1)
    bound_tflops = min(2500.0, flops / nbytes * 8.0)
    return {"batch_states": batch, "train_states_per_s": batch / t_train, "infer_states_per_s": batch / t_infer,
            "gemm_layer1": {"M": M, "N": H1, "K": PER_TEST, "us": t_gemm * 1e6,
                            "tflops": flops / t_gemm / 1e12, "peak_tflops": 2500.0,
                            "frac": flops / t_gemm / 2.5e15,
                            "roofline": {"bound": "hbm", "flop_per_byte": flops / nbytes,
                                         "algorithmic_bytes": nbytes, "achieved_gbs": nbytes / t_gemm / 1e9,
                                         "peak_gbs": 8000.0, "bound_tflops": bound_tflops,
                                         "frac": nbytes / t_gemm / 8e12}},
            "params": net.n_params, "timing": "host clock around device-synchronised repetitions"}


def wavefront_bench(task, D=8, s=3):
    """The GPU wavefront producer (exact-OEP level expansion, the expand_*
    kernels) of the bench task down to split s of depth D: device time from
    the handle's HIP events, children and survivors per level summed, and
    the algorithmic HBM bytes of the table and frontier (per child a 16-byte
    slot probed in each of two passes, an 8-byte atomicMin and its 8-byte
    key written and read back; per survivor its 18-byte record) against the
    8 TB/s peak."""
    from pca_amd import Solver
    from pca_amd._lib import lib
    with Solver(task) as sv:
        lib().pcagpu_frontier(sv._h, D, s, None, 0)   # first call: buffers
        sv.reset_stats()
        n = lib().pcagpu_frontier(sv._h, D, s, None, 0)
        st = sv.stats()
    secs = st["expand_ns"] * 1e-9
    children, kept = st["frontier_nodes"], st["frontier_kept"]
    nbytes = children * 56 + kept * 18
    return {"workload": f"is_power2_32 wavefront, depth {D}, levels 0..{s}", "prefixes": n, "children": children,
            "kept": kept, "expand_ms": secs * 1e3, "children_per_s": children / secs,
            "hbm": {"algorithmic_bytes": nbytes, "achieved_gbs": nbytes / secs / 1e9, "peak_gbs": 8000.0,
                    "frac": nbytes / secs / 8e12}}


def completion_bench(reps=3):
    """Leaf completion (search_exhaustive_complete) on the GPU: the
    two-input fixture states with a budget of 3 instructions (the reference
    loop's completion radius), per-call time against the reference
    library's one-core time recorded with the fixtures."""
    import statistics
    from pca_amd import Task, Solver, search_state
    cdir = os.path.join(ROOT, "tests", "golden", "complete")
    with open(os.path.join(cdir, "cases.json")) as f:
        cases = [c for c in json.load(f)["cases"] if c["kind"] == "heavy"]
    out = []
    for c in cases:
        task = Task.load(os.path.join(cdir, "tasks", c["task"] + ".json"))
        st = search_state.after(task, c["path"])
        with Solver(task) as sv:
            sv.exhaustive_complete(st, c["max_depth"])   # first call: tables, buffers
            ts = []
            for _ in range(reps):
                t0 = time.perf_counter()
                found, n, ins = sv.exhaustive_complete(st, c["max_depth"])
                ts.append(time.perf_counter() - t0)
            explored = sv.exhaustive_states_explored()
        assert (found, ins) == (c["found"], c["insns"]), c["task"]
        ms = statistics.median(ts) * 1e3
        out.append({"task": c["task"], "path": c["path"], "max_depth": c["max_depth"], "found": found,
                    "gpu_ms": ms, "gpu_children": explored, "reference_s": c["seconds"],
                    "reference_explored": c["explored"], "speedup": c["seconds"] / (ms / 1e3)})
    res = {"workload": "two-input tasks, start states 2 instructions deep, budget 3 (R = 3)",
           "cases": out, "gpu_ms_total": sum(x["gpu_ms"] for x in out),
           "reference_s_total": sum(x["reference_s"] for x in out),
           "reference_timing": "reference search_exhaustive_complete, one core of the build host"}
    res["speedup_total"] = res["reference_s_total"] / (res["gpu_ms_total"] / 1e3)
    return res


def completion_cpu_baseline():
    """The oracle's restatement of search_exhaustive_complete (32-bit OEP
    table as in the reference, one thread) on this host, on the two-input
    budget-3 fixture state the reference finishes fastest."""
    from oracle import oracle as O
    cdir = os.path.join(ROOT, "tests", "golden", "complete")
    with open(os.path.join(cdir, "cases.json")) as f:
        cases = [c for c in json.load(f)["cases"] if c["kind"] == "heavy"]
    c = min(cases, key=lambda x: x["seconds"])
    ot = O.load_task(os.path.join(cdir, "tasks", c["task"] + ".json"))
    t0 = time.perf_counter()
    O.complete(ot, O.state_after(ot, c["path"]), c["max_depth"], O.OEP_REF32)
    return {"task": c["task"], "path": c["path"], "max_depth": c["max_depth"], "seconds": time.perf_counter() - t0,
            "cores": 1, "kind": "port"}


def window(task, D, s, B, warmup, steps, progress=False):
    """The fixed-depth window: DFS-ordered split-s prefixes of the depth-D
    tree (host DFS, exact OEP), B per step, each step one flush through the
    scan kernel (prefixes resident before the timed region).  Returns the
    wall time of the timed steps, the handle's counters over them and the
    device name."""
    import ctypes as C
    import pca_amd
    need = B * (warmup + steps)
    prefixes, _, _ = pca_amd.host_frontier(task, D, s, cap=need)
    if len(prefixes) < need:
        raise RuntimeError("not enough prefixes")
    with pca_amd.Solver(task) as sv:
        lib = pca_amd.lib()
        arr = (pca_amd.Prefix * need)()
        for i, (ins, live) in enumerate(prefixes):
            for q, v in enumerate(ins):
                arr[i].insns[q] = v
            arr[i].num_insns = s
            arr[i].live = live

        def step(i):
            rc = lib.pcagpu_begin_depth(sv._h, D, s)
            assert rc == 0, rc
            ptr = C.cast(C.addressof(arr) + i * B * C.sizeof(pca_amd.Prefix), C.POINTER(pca_amd.Prefix))
            rc = lib.pcagpu_submit(sv._h, ptr, B, i * B)
            assert rc == 0, rc
            rc = lib.pcagpu_flush(sv._h)
            assert rc >= 0, rc
            n = lib.pcagpu_end_depth(sv._h, None, 0)
            assert n >= 0, n

        for i in range(warmup):
            step(i)
            if progress:
                print(f"warmup step {i} done", file=sys.stderr, flush=True)
        sv.reset_stats()
        t0 = time.perf_counter()
        for i in range(warmup, warmup + steps):
            step(i)
            if progress:
                print(f"step {i - warmup} done", file=sys.stderr, flush=True)
        t1 = time.perf_counter()
        return t1 - t0, sv.stats(), sv.device_name


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    # the first launches try the scan tiles (one each) before the winner is kept
    ap.add_argument("--warmup", type=int, default=6)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--split", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--suite", type=int, default=1000, help="config-3 suite size (0: skip)")
    ap.add_argument("--no-value", action="store_true")
    ap.add_argument("--no-complete", action="store_true")
    ap.add_argument("--suite-budget", type=int, default=3)
    ap.add_argument("--suite-deep-budget", type=int, default=4,
                    help="a second suite field at this depth budget (0: skip)")
    ap.add_argument("--serving-requests", type=int, default=200)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the search path shards by prefix batches, see DESIGN.md)", file=sys.stderr)
        return 2

    import pca_amd
    from pca_amd import taskgen
    task = taskgen.is_power2_32(42)
    task_path = os.path.join(ROOT, "tests", "golden", "tasks", "is_power2_32.json")
    D, s, B = args.depth, args.split, args.batch
    wall, st, dev = window(task, D, s, B, args.warmup, args.steps, progress=True)
    cand = st["candidates"]
    launches = max(st["launches"], 1)
    kernel_s = st["kernel_ns"] * 1e-9
    avg_launch = kernel_s / launches
    ops = OPS_PER_TEST * st["tests_evaluated"] + OPS_PER_PARENT_TEST * task.num_tests * st["parents"]
    achieved = ops / launches / avg_launch
    line = {
        "metric": "candidates/s (1 GPU)",
        "value": cand / wall,
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic (seeded is_power2_32 tests, DFS-ordered depth-5 prefixes)",
        "config": {"workload": f"is_power2_32 hybrid IDDFS window: kernel depth {D}, split {s}, "
                               f"{D - s}-insn suffix, {B} prefixes per step",
                   "tests": task.num_tests, "kernel_depth": D, "split": s, "prefixes_per_step": B,
                   "device": dev},
        "roofline": {"bound": "valu-int", "achieved": achieved / 1e12, "peak": PEAK_INT_OPS / 1e12,
                     "unit": "Tint-ops/s", "frac": achieved / PEAK_INT_OPS, "traffic": None},
        # the per-(prefix, suffix) figure of the reference formulation: every
        # suffix replayed whole (k VM steps of 8 ops) plus 7 register compares
        # per test it reaches; counting one test per suffix is a lower bound.
        # It credits work this kernel never does (shared suffix prefixes,
        # value groups), so it can pass the peak; the figure above is the
        # kernel's own work.
        "roofline_per_suffix": {"ops_per_suffix_test": (D - s) * 8 + 7,
                                "achieved_lower_bound": st["suffixes"] * ((D - s) * 8 + 7) / kernel_s / 1e12,
                                "frac_lower_bound": st["suffixes"] * ((D - s) * 8 + 7) / kernel_s / PEAK_INT_OPS,
                                "unit": "Tint-ops/s"},
        "kernel": {"name": "scan_kernel", "avg_launch_ms": avg_launch * 1e3, "launches": launches,
                   "candidates": cand, "suffixes": st["suffixes"], "tests_evaluated": st["tests_evaluated"],
                   "parents": st["parents"], "raw_hits": st["raw_hits"],
                   "kernel_candidates_per_s": cand / kernel_s if kernel_s else None,
                   "tile": {"block_kids": st["scan_tile"] & 0xFF, "mask_tests": st["scan_tile"] >> 8}},
    }
    # HBM bytes per launch from the committed rocprofv3 PMC passes of this
    # kernel and workload (FETCH_SIZE x 2.0 + WRITE_SIZE x 1.0, copy-calibrated)
    prof = os.path.join(ROOT, "profiles", "scan_kernel_is_power2_32_d8s5.json")
    if os.path.exists(prof) and (D, s, B) == (8, 5, 16):
        with open(prof) as f:
            line["roofline"]["traffic"] = json.load(f)["traffic_bytes_per_launch"]
    if not args.no_serving:
        line["serving"] = serving_bench(n_req=args.serving_requests)
    if args.suite:
        line["suite"] = suite_bench(args.suite, args.suite_budget)
        if args.suite_deep_budget:
            # the deeper budget: the tasks whose optimum needs it dominate
            line["suite_deep"] = suite_bench(args.suite, args.suite_deep_budget)
    if not args.no_value:
        line["value_net"] = value_bench()
        print("value net done", file=sys.stderr, flush=True)
    line["wavefront"] = wavefront_bench(task)
    print("wavefront done", file=sys.stderr, flush=True)
    if not args.no_complete:
        line["leaf_completion"] = completion_bench()
        print("leaf completion done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        if not args.no_complete:
            line["leaf_completion"]["cpu_baseline"] = completion_cpu_baseline()
        cb = cpu_baseline(task_path, args.cpu_seconds)
        line["cpu_baseline"] = cb
        if cb:
            line["speedup_vs_cpu"] = line["value"] / cb["value"]
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
