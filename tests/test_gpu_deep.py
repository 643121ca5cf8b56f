"""Parity at the depths and test counts of the workloads: the published
optima (max 7, saturating_add 8 and is_power2 7 instructions), config 2's
32-test is_power2 at its proven optimum, generated 26-32-test tasks with
solutions at every split, sliced differentials at kernel depth 8 with hits,
and k = 0 leaf scans.

The expected sets are the oracle's (tests/golden/exact/, written in the
build container by tests/golden/make_golden.py: per depth the exact-key DFS
to D-1 and the last position brute force on the VM, pinned against the
oracle's full IDDFS by the CPU tests), or the oracle run here on slices.
"""
import json
import os
import random

import pytest

from conftest import GOLDEN, golden_task_path

pytestmark = pytest.mark.gpu

DEEP = [("max", 3), ("is_power2", 4), ("is_power2_32", 4), ("saturating_add", 4)]
GEN = sorted((f[:-5].rsplit("_d", 1)[0], int(f[:-5].rsplit("_d", 1)[1]))
             for f in os.listdir(os.path.join(GOLDEN, "exact")) if f.startswith("gen_"))


def _fixture(name, d):
    with open(os.path.join(GOLDEN, "exact", f"{name}_d{d}.json")) as f:
        return json.load(f)


def _check(res, want):
    assert res.optimal_depth == want["optimal_depth"]
    assert res.num_solutions == want["num_solutions"]
    assert [s[0] for s in res.solutions] == want["solutions"]
    assert all(s[4] for s in res.solutions)


@pytest.mark.parametrize("name,maxd", DEEP)
def test_published_optimum_matches_oracle(pca, name, maxd):
    """IDDFS to the optimum with the automatic split: the same optimal
    length as the reference's published programs and the oracle's exact
    canonical set, every program re-verified on the CPU VM."""
    want = _fixture(name, maxd)
    t = pca.Task.load(golden_task_path(name))
    print(f"[deep] {name} d{maxd} tests={t.num_tests}", flush=True)
    with pca.Solver(t) as sv:
        res = sv.solve(maxd, cap=1 << 16)
    _check(res, want)


@pytest.mark.parametrize("name,maxd", GEN)
@pytest.mark.parametrize("split", [-1, 0, 1, 2])
def test_generated_32_test_tasks_match_oracle(pca, name, maxd, split):
    """Tasks of 26-32 tests with solutions: every split gives the oracle's
    set -- a 1- to 3-instruction suffix brute-forced on the device, hits
    emitted over all tests, and the long suffixes cut to the exact OEP set."""
    want = _fixture(name, maxd)
    t = pca.Task.load(golden_task_path(name))
    print(f"[gen] {name} d{maxd} split={split} tests={t.num_tests}", flush=True)
    with pca.Solver(t) as sv:
        res = sv.solve(maxd, split=split, cap=1 << 16)
        st = sv.stats()
    _check(res, want)
    assert st["raw_hits"] >= want["num_solutions"] > 0


@pytest.mark.parametrize("name,maxd", GEN[:3])
def test_generated_32_test_tasks_host_producer(pca, name, maxd):
    want = _fixture(name, maxd)
    t = pca.Task.load(golden_task_path(name))
    print(f"[gen-host] {name} d{maxd}", flush=True)
    with pca.Solver(t) as sv:
        res = sv.solve(maxd, producer="host", cap=1 << 16)
    _check(res, want)


# ---------------------------------------------------------------- slices
def _planted_prefixes(t, kernel, body, n_fill, variants):
    """Prefixes made of n_fill LDI fillers on registers the solution
    `kernel` never names, then `body` (a head of the kernel).  Fillers do
    not touch flags, so the kernel runs as it does alone."""
    used = set()
    for c in kernel:
        op = (c >> 11) & 31
        used.add((c >> 8) & 7)
        if op not in (16, 18, 19):            # LDI, ADDI, BR read no rs/rt
            used.update(((c >> 5) & 7, (c >> 2) & 7))
    free = [r for r in range(1, 8) if r not in used]
    assert free, "no register left for fillers"
    rng = random.Random(17)
    out = []
    for v in range(variants):
        fill = [(16 << 11) | (free[(v + q) % len(free)] << 8) | rng.choice([0, 1, 2, 3, 255]) for q in range(n_fill)]
        ins = fill + list(body)
        live = sum(1 << (q + 1) for q in range(t.num_inputs))
        for c in ins:
            if ((c >> 11) & 31) not in (13, 19):
                live |= 1 << ((c >> 8) & 7)
        out.append((tuple(ins), live & 0xFE))
    return out


def _random_prefixes(pca, t, D, s, n, seed):
    rng = random.Random(seed)
    out = []
    for _ in range(n):
        live = sum(1 << (q + 1) for q in range(t.num_inputs))
        ins = []
        for pos in range(s):
            c = rng.choice(pca.gen_candidates(live, pos, D))
            ins.append(c)
            if ((c >> 11) & 31) not in (13, 19):
                live |= 1 << ((c >> 8) & 7)
        out.append((tuple(ins), live & 0xFE))
    return out


def _kernel_of(prog, n_in, D):
    return prog[n_in:n_in + D]


def test_is_power2_32_depth8_last_position_slice(pca, oracle):
    """Config 2's task at kernel depth 8, split 7 (k = 1): the device's raw
    hits below 7-instruction prefixes -- random DFS paths and prefixes that
    end in a proven 4-instruction solution's first three instructions --
    equal the oracle's VM brute force of the same children (register mask
    over r1..r7 and can-never-pass flag, in canonical order)."""
    t = pca.Task.load(golden_task_path("is_power2_32"))
    ot = oracle.task_from_bytes(bytes(t))
    fx = _fixture("is_power2_32", 4)
    kernels = [_kernel_of(p, t.num_inputs, 4) for p in fx["solutions"][::9]]
    prefixes = _random_prefixes(pca, t, 8, 7, 24, seed=5)
    for k in kernels:
        # the solution's last instruction lands at position 7, its forward
        # branches at most on the OUT, as in the 4-instruction program
        prefixes += _planted_prefixes(t, k, k[:-1], 4, 2)
    print(f"[slice k=1] is_power2_32 D=8 s=7 prefixes={len(prefixes)}", flush=True)
    want, _ = oracle.last_level(ot, 8, prefixes=prefixes)
    with pca.Solver(t) as sv:
        raw = sv.raw_depth(8, 7, prefixes)
    got = [(p, js, mask, flags & 1) for p, flags, mask, js in raw]
    assert got == [(p, (j,), mask, dead) for p, j, mask, dead, _ in want]
    assert sum(1 for r in want if not r[3]) >= len(kernels)


def test_is_power2_32_depth8_two_position_slice(pca, oracle):
    """The same at split 6 (k = 2) below a planted prefix -- three fillers
    and a proven solution's first three instructions, so its last one is
    among the position-6 candidates.  The oracle enumerates the two-position
    suffix space as the last position below every position-6 child (the
    VM brute force of or_last_level); every raw row, mask and flag equal."""
    t = pca.Task.load(golden_task_path("is_power2_32"))
    ot = oracle.task_from_bytes(bytes(t))
    fx = _fixture("is_power2_32", 4)
    kernels = [_kernel_of(p, t.num_inputs, 4) for p in fx["solutions"]]
    plain = [k for k in kernels if all(((c >> 11) & 31) != 19 for c in k)]
    kernel = (plain or kernels)[0]
    prefixes = _planted_prefixes(t, kernel, kernel[:-1], 3, 1)
    ins, live = prefixes[0]
    print(f"[slice k=2] is_power2_32 D=8 s=6 prefix={ins}", flush=True)
    L = oracle.lib()
    children = [(tuple(ins) + (c,), L.or_live_after(live, c)) for c in oracle.gen_candidates(live, 6, 8)]
    rows, _ = oracle.last_level(ot, 8, prefixes=children)
    want = [(0, (ci, j), mask, dead) for ci, j, mask, dead, _ in rows]
    with pca.Solver(t) as sv:
        raw = sv.raw_depth(8, 6, prefixes)
    got = [(p, js, mask, flags & 1) for p, flags, mask, js in raw]
    assert got == want
    assert any(not r[3] for r in want)


# ------------------------------------------------------------ k = 0 leaves
@pytest.mark.parametrize("name,maxd", [("negate", 3), ("double", 3), ("add", 3), ("gen_chain_036", 2),
                                       ("gen_chain_040", 2)])
def test_split_equal_to_depth_leaf_scan_matches_fixture(pca, name, maxd):
    """split = D: the wavefront expands every level and the scan only checks
    the finished programs' registers (k = 0, gpu/wavefront.cu:421-468)."""
    want = _fixture(name, maxd)
    t = pca.Task.load(golden_task_path(name))
    print(f"[k=0] {name} d{maxd}", flush=True)
    with pca.Solver(t) as sv:
        res = sv.solve(maxd, split=8, cap=1 << 16)
    _check(res, want)


@pytest.mark.parametrize("name,D", [("ident0", 2), ("add", 2), ("gen_chain_040", 2)])
def test_leaf_scan_raw_hits_match_oracle(pca, oracle, name, D):
    """Raw k = 0 hits below the oracle's level-D frontier: each program
    whose replay state has registers equal to the target on every test,
    with that mask and whether some test cannot halt."""
    t = pca.Task.load(golden_task_path(name))
    ot = oracle.task_from_bytes(bytes(t))
    front = oracle.frontier(ot, D, D)
    # programs with hits (the first visitors among the oracle's last-level
    # rows) and random others
    par = oracle.frontier(ot, D, D - 1)
    lrows, _ = oracle.last_level(ot, D, prefixes=par)
    owners = {tuple(par[p][0]) + (oracle.gen_candidates(par[p][1], D - 1, D)[j],)
              for p, j, _, _, first in lrows if first}
    own = [i for i, (ins, _) in enumerate(front) if tuple(ins) in owners]
    rng = random.Random(9)
    idxs = sorted(set(own[:150]) | set(rng.sample(range(len(front)), min(150, len(front)))))
    sub = [front[i] for i in idxs]
    T = ot.num_tests
    tgt = [ot.tests[k].out[0] for k in range(T)]
    want = []
    for i, (ins, live) in enumerate(sub):
        st = oracle.replay_state(ot, list(ins))
        mask = sum(1 << r for r in range(1, 8) if all(st[k][r] == tgt[k] for k in range(T)))
        if not mask:
            continue
        prog = [(23 << 11) | ((q + 1) << 8) | ot.input_ports[q] for q in range(ot.num_inputs)]
        prog += list(ins) + [(24 << 11) | (1 << 8) | ot.output_ports[0], 1 << 11]
        # can never pass: some test does not halt within the cycle limit
        dead = int(any(_halts(oracle, ot, prog, k) is False for k in range(T)))
        want.append((i, (), mask, dead))
    print(f"[k=0 raw] {name} D={D} prefixes={len(sub)} rows={len(want)}", flush=True)
    with pca.Solver(t) as sv:
        raw = sv.raw_depth(D, D, sub)
    got = [(p, js, mask, flags & 1) for p, flags, mask, js in raw]
    assert got == want
    assert want


def _halts(oracle, ot, prog, k):
    one = oracle.parse_task(json.dumps({
        "name": "one", "max_cycles": int(ot.max_cycles),
        "input_ports": [ot.input_ports[q] for q in range(ot.num_inputs)], "output_ports": [ot.output_ports[0]],
        "tests": [{"in": [ot.tests[k].inp[q] for q in range(ot.num_inputs)], "out": [0]}]}))
    # the program halts iff it passes with the output it writes
    st = oracle.replay_state(ot, prog[ot.num_inputs:-2])
    one.tests[0].out[0] = st[k][1]
    return oracle.evaluate(one, prog) == 1


def test_server_deep_requests_match_oracle(pca):
    """Deep requests through the continuous-batching server (device-resident
    frontiers of millions of prefixes, scanned in place, next to small
    requests): the same sets as the oracle's fixtures."""
    from pca_amd.serving import Server, DONE
    reqs = [("is_power2_32", 4), ("max", 3), ("is_power2", 4), ("gen_chain_016", 3), ("negate", 3)]
    print("[server-deep]", reqs, flush=True)
    with Server(max_slots=4) as sv:
        ids = [sv.submit(pca.Task.load(golden_task_path(n)), d) for n, d in reqs]
        recs = {r["id"]: r for r in sv.drain()}
        for rid, (name, d) in zip(ids, reqs):
            assert recs[rid]["status"] == DONE
            sols, n = sv.solutions(rid)
            if name == "negate":
                assert recs[rid]["optimal_depth"] == 1 and n == 7
                continue
            want = _fixture(name, d)
            assert (recs[rid]["optimal_depth"], n) == (want["optimal_depth"], want["num_solutions"])
            assert [x[0] for x in sols] == want["solutions"][:4096]
