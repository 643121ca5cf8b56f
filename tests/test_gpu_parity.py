"""Parity of the HIP path (through the C ABI) against the CPU oracle.

Every comparison is bit-exact: frontiers, raw kernel hits, and the final
canonically ordered program sets.
"""
import random

import pytest

from conftest import golden_task_path

pytestmark = pytest.mark.gpu

TASKS = ["negate", "double", "square", "add", "abs", "relu", "sign", "max"]


def _pair(pca, oracle, name):
    t = pca.Task.load(golden_task_path(name))
    return t, oracle.task_from_bytes(bytes(t))


@pytest.mark.parametrize("name,D,s", [("abs", 3, 2), ("abs", 3, 1), ("max", 3, 2), ("add", 2, 1),
                                      ("relu", 3, 2), ("sign", 2, 2), ("negate", 3, 2), ("negate", 4, 2)])
def test_wavefront_frontier_matches_oracle(pca, oracle, name, D, s):
    t, ot = _pair(pca, oracle, name)
    print(f"[frontier] {name} D={D} s={s}", flush=True)
    want = oracle.frontier(ot, D, s)
    with pca.Solver(t) as sv:
        got = sv.frontier(D, s)
    assert got == want


def _oracle_suffix_rows(oracle, ot, D, s, prefixes, idxs):
    rows = []
    for i in idxs:
        ins, live = prefixes[i]
        hits, n, _ = oracle.suffix_hits(ot, D, list(ins) + [0] * (8 - len(ins)), s, live)
        assert n == len(hits)
        for h in hits:
            rows.append((i, tuple(h[:-1]), h[-1]))
    return rows


def _gpu_suffix_rows(pca, sv, D, s, prefixes, idxs):
    sub = [prefixes[i] for i in idxs]
    raw = sv.raw_depth(D, s, sub)
    rows = []
    for pidx, flags, mask, js in raw:
        if flags & 1:
            continue
        ins, live = sub[pidx]
        # live after the suffix, from the grammar order
        lv = live
        for q, j in enumerate(js):
            c = pca.gen_candidates(lv, s + q, D)[j]
            op, rd = (c >> 11) & 31, (c >> 8) & 7
            if op not in (13, 19) and rd != 0:
                lv |= 1 << rd
        for r in range(1, 8):
            if mask >> r & 1 and lv >> r & 1:
                rows.append((idxs[pidx], js, r))
    return rows


@pytest.mark.parametrize("name,D", [("abs", 3), ("max", 3), ("relu", 3), ("add", 2), ("negate", 3),
                                    ("sign", 3)])
def test_scan_k1_raw_hits_match_bruteforce(pca, oracle, name, D):
    t, ot = _pair(pca, oracle, name)
    s = D - 1
    prefixes = oracle.frontier(ot, D, s)
    rng = random.Random(7)
    idxs = sorted(set(list(range(min(40, len(prefixes)))) + rng.sample(range(len(prefixes)),
                                                                       min(60, len(prefixes)))))
    print(f"[scan k=1] {name} D={D} prefixes={len(idxs)}", flush=True)
    want = _oracle_suffix_rows(oracle, ot, D, s, prefixes, idxs)
    with pca.Solver(t) as sv:
        got = _gpu_suffix_rows(pca, sv, D, s, prefixes, idxs)
    assert got == want


@pytest.mark.parametrize("name,D,s", [("abs", 3, 1), ("relu", 3, 1)])
def test_scan_k2_raw_hits_match_bruteforce(pca, oracle, name, D, s):
    t, ot = _pair(pca, oracle, name)
    prefixes = oracle.frontier(ot, D, s)
    rng = random.Random(11)
    idxs = sorted(rng.sample(range(len(prefixes)), min(2, len(prefixes))))
    print(f"[scan k=2] {name} D={D} s={s} prefixes={idxs}", flush=True)
    want = _oracle_suffix_rows(oracle, ot, D, s, prefixes, idxs)
    with pca.Solver(t) as sv:
        got = _gpu_suffix_rows(pca, sv, D, s, prefixes, idxs)
    assert got == want


def _exact_fixture(name, maxd):
    import json
    import os
    from conftest import GOLDEN
    path = os.path.join(GOLDEN, "exact", f"{name}_d{maxd}.json")
    with open(path) as f:
        return json.load(f)


@pytest.mark.parametrize("name,maxd", [("negate", 3), ("double", 3), ("square", 3), ("add", 3)])
@pytest.mark.parametrize("producer", ["gpu", "host"])
def test_solve_matches_oracle_live(pca, oracle, name, maxd, producer):
    t, ot = _pair(pca, oracle, name)
    print(f"[solve] {name} producer={producer}", flush=True)
    want = oracle.solve(ot, maxd, oracle.OEP_EXACT, max_solutions=0)
    with pca.Solver(t) as sv:
        res = sv.solve(maxd, producer=producer, cap=1 << 16)
    assert res.optimal_depth == want["optimal_depth"]
    assert res.num_solutions == want["num_solutions"]
    assert [s[0] for s in res.solutions] == want["solutions"]
    assert all(s[4] for s in res.solutions)


@pytest.mark.parametrize("name,maxd", [("abs", 3), ("relu", 3), ("sign", 3), ("clamp", 2), ("popcount", 2),
                                       ("is_power2", 2), ("max", 2), ("negate_16", 5), ("is_power2_32", 2)])
@pytest.mark.parametrize("producer", ["gpu", "host"])
def test_solve_matches_exact_fixture(pca, name, maxd, producer):
    want = _exact_fixture(name, maxd)
    t = pca.Task.load(golden_task_path(name))
    print(f"[solve-fixture] {name} d{maxd} producer={producer}", flush=True)
    with pca.Solver(t) as sv:
        res = sv.solve(maxd, producer=producer, cap=1 << 16)
    assert res.optimal_depth == want["optimal_depth"]
    assert res.num_solutions == want["num_solutions"]
    assert [s[0] for s in res.solutions] == want["solutions"]
    assert all(s[4] for s in res.solutions)


def _oracle_all_raw(pca, oracle, ot, D, s, prefixes, idxs):
    """Every child (k = 1) whose final state has a register r1..r7 equal to
    the target on all tests, with that mask and whether the program cannot
    pass (a test faults or times out) -- the device hit semantics."""
    T = ot.num_tests
    tgt = [ot.tests[t].out[0] for t in range(T)]
    rows = []
    for bi, i in enumerate(idxs):
        ins, live = prefixes[i]
        cands = pca.gen_candidates(live, s, D)
        for j, c in enumerate(cands):
            kern = list(ins) + [c]
            st = oracle.replay_state(ot, kern)
            mask = 0
            for r in range(1, 8):
                if all(st[t][r] == tgt[t] for t in range(T)):
                    mask |= 1 << r
            if not mask:
                continue
            prog = [(23 << 11) | ((q + 1) << 8) | ot.input_ports[q] for q in range(ot.num_inputs)]
            prog += kern + [(24 << 11) | (1 << 8) | ot.output_ports[0], 1 << 11]
            # dead: some test does not halt (fault / cycle limit) whatever is output
            dead = 0
            for t in range(T):
                single = oracle.OrTask()
                import ctypes as C
                C.memmove(C.byref(single), C.byref(ot), C.sizeof(single))
                single.tests[0] = ot.tests[t]
                single.num_tests = 1
                single.tests[0].out[0] = st[t][1]
                if oracle.evaluate(single, prog) != 1:
                    dead = 1
                    break
            rows.append((bi, (j,), mask, dead))
    return rows


@pytest.mark.parametrize("name,D", [("ident0", 2), ("ident0", 3), ("is_power2", 3), ("abs", 3)])
def test_scan_raw_hits_with_shadow_hits(pca, oracle, name, D):
    t, ot = _pair(pca, oracle, name)
    s = D - 1
    prefixes = oracle.frontier(ot, D, s)
    rng = random.Random(5)
    idxs = sorted(rng.sample(range(len(prefixes)), min(40, len(prefixes))))
    print(f"[raw+shadow] {name} D={D} prefixes={len(idxs)}", flush=True)
    want = _oracle_all_raw(pca, oracle, ot, D, s, prefixes, idxs)
    with pca.Solver(t) as sv:
        raw = sv.raw_depth(D, s, [prefixes[i] for i in idxs])
    got = [(p, js, mask, flags & 1) for p, flags, mask, js in raw]
    assert got == want


@pytest.mark.parametrize("name,maxd,split", [("abs", 3, 1), ("abs", 3, 0), ("relu", 3, 1), ("relu", 3, 0),
                                             ("sign", 3, 1), ("negate_16", 5, 0), ("is_power2_32", 2, 0),
                                             ("max", 2, 0)])
def test_solve_any_split_matches_exact_fixture(pca, name, maxd, split):
    """Suffixes longer than one instruction are brute-forced on the device
    and then cut down to the exact OEP set (state-match passes + host
    resolution): the result must not depend on the split."""
    want = _exact_fixture(name, maxd)
    t = pca.Task.load(golden_task_path(name))
    print(f"[solve-split] {name} d{maxd} split={split}", flush=True)
    with pca.Solver(t) as sv:
        res = sv.solve(maxd, split=split, cap=1 << 16)
    assert res.optimal_depth == want["optimal_depth"]
    assert res.num_solutions == want["num_solutions"]
    assert [s[0] for s in res.solutions] == want["solutions"]


@pytest.mark.parametrize("name,maxd,split", [("abs", 3, 1), ("is_power2_32", 2, 1), ("sign", 3, 1)])
def test_adaptive_test_order_keeps_results(pca, name, maxd, split):
    """Small batches make many scan launches, and the device test order is
    re-sorted after each one: the solution set must equal the fixed-order
    run and the exact fixture."""
    want = _exact_fixture(name, maxd)
    t = pca.Task.load(golden_task_path(name))
    print(f"[adaptive-order] {name} d{maxd} split={split}", flush=True)
    got = {}
    for fixed in (True, False):
        with pca.Solver(t, batch_capacity=4, fixed_test_order=fixed) as sv:
            res = sv.solve(maxd, split=split, cap=1 << 16)
            got[fixed] = ([s[0] for s in res.solutions], res.optimal_depth, res.num_solutions,
                          sv.stats()["launches"], sv.stats()["tests_evaluated"])
    print("fixed", got[True][3:], "adaptive", got[False][3:], flush=True)
    assert got[True][:3] == got[False][:3]
    assert got[False][:3] == (want["solutions"], want["optimal_depth"], want["num_solutions"])
    assert got[False][3] > 4


@pytest.mark.parametrize("name,D,s", [("abs", 3, 1), ("relu", 3, 1), ("sign", 3, 2)])
def test_graph_flush_replays_match_eager(pca, oracle, name, D, s):
    """Flushes of 3 prefixes replay the scan graph many times, with the
    library's own eager copies and syncs between replays; the raw hits of
    every replay, in order, must equal plain stream launches' hits."""
    t = pca.Task.load(golden_task_path(name))
    prefixes, _, _ = pca.host_frontier(t, D, s)
    print(f"[graph] {name} D{D} s{s} prefixes={len(prefixes)}", flush=True)
    rows = {}
    for eager in (True, False):
        with pca.Solver(t, batch_capacity=3, eager_flush=eager, fixed_test_order=True, scan_tile=(64, 2)) as sv:
            rows[eager] = sv.raw_depth(D, s, prefixes)
            launches = sv.stats()["launches"]
    assert launches >= len(prefixes) // 3
    assert rows[False] == rows[True]
    assert len(rows[True]) > 0


@pytest.mark.parametrize("tile", [(64, 1), (64, 2), (64, 3), (32, 2), (16, 3), (1, 2)])
@pytest.mark.parametrize("name,maxd,split", [("abs", 3, 0), ("sign", 3, 1), ("is_power2_32", 2, 0)])
def test_scan_tiles_keep_results(pca, name, maxd, split, tile):
    """Every scan tile (children per block, tests run by wave masks) the
    autotuner may pick gives the exact fixture's solution set."""
    want = _exact_fixture(name, maxd)
    t = pca.Task.load(golden_task_path(name))
    print(f"[tile] {name} d{maxd} split={split} tile={tile}", flush=True)
    with pca.Solver(t, scan_tile=tile) as sv:
        res = sv.solve(maxd, split=split, cap=1 << 16)
        assert sv.stats()["scan_tile"] == tile[0] | (tile[1] << 8)
    assert res.optimal_depth == want["optimal_depth"]
    assert res.num_solutions == want["num_solutions"]
    assert [s[0] for s in res.solutions] == want["solutions"]


@pytest.mark.parametrize("opname", ["ADD", "SUB", "MUL", "DIV", "MOD", "AND", "OR", "XOR", "SHL", "SHR",
                                    "ASR", "CMP"])
def test_device_alu_exhaustive_matches_oracle(pca, opname):
    """Every (a, b) in 2^16 x 2^16 on the device vs the oracle's digest."""
    import json
    import os
    from conftest import GOLDEN
    from pca_amd._lib import alu_digest
    with open(os.path.join(GOLDEN, "alu_digests.json")) as f:
        g = json.load(f)[opname]
    t = pca.Task.load(golden_task_path("negate"))
    with pca.Solver(t) as sv:
        assert str(alu_digest(g["op"], 0, 65536, solver=sv)) == g["full"]


def test_cli_hybrid_drop_in(pca):
    import os
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "pca_amd", "gpu_hybrid")
    out = subprocess.run([exe, golden_task_path("abs"), "4", "5"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "Optimal solution length: 6 instructions (kernel depth 3)" in out.stdout
    assert out.stdout.count("cpu_verify=8/8 OK") == 5
    for exe_name in ("gpu_wavefront", "gpu_enumerate_oep"):
        o2 = subprocess.run([os.path.join(ROOT, "pca_amd", exe_name), golden_task_path("abs"), "4", "5"],
                            capture_output=True, text=True, timeout=120)
        assert o2.returncode == 0, o2.stderr
        # the same canonical programs, whichever producer made the prefixes
        assert o2.stdout.split("Optimal solution length")[1].split("Total:")[0] == \
            out.stdout.split("Optimal solution length")[1].split("Total:")[0]
