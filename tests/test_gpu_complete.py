"""Leaf completion on the GPU (pcagpu_complete) against the reference
library's answers (tests/golden/complete/cases.json) and the oracle's exact
restatement of search_exhaustive_complete (src/search_state.c:339-445)."""
import json
import os
import random
import time

import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu

CDIR = os.path.join(GOLDEN, "complete")
with open(os.path.join(CDIR, "cases.json")) as _f:
    CASES = json.load(_f)["cases"]
TASKS = sorted({c["task"] for c in CASES})


def _cid(c):
    return f"{c['task']}-{c['kind']}-{'.'.join(map(str, c['path']))}-{c['max_depth']}"


@pytest.fixture(scope="module")
def solvers(pca):
    out = {}
    for name in TASKS:
        t = pca.Task.load(os.path.join(CDIR, "tasks", name + ".json"))
        out[name] = (t, pca.Solver(t))
    yield out
    for _, sv in out.values():
        sv.close()


@pytest.mark.parametrize("c", CASES, ids=_cid)
def test_completion_matches_reference(pca, solvers, c):
    """Same found flag and the same instructions as the reference's
    search_exhaustive_complete, every fixture case (the two-input cases with
    a budget of 3 take the reference 8-90 s on one core)."""
    t, sv = solvers[c["task"]]
    st = pca.search_state.after(t, c["path"])
    print(f"[complete] {_cid(c)}", flush=True)
    t0 = time.perf_counter()
    found, n, ins = sv.exhaustive_complete(st, c["max_depth"])
    dt = time.perf_counter() - t0
    if c["kind"] == "heavy":
        print(f"  gpu {dt * 1e3:.1f} ms, reference {c['seconds']:.2f} s, explored {sv.exhaustive_states_explored()}",
              flush=True)
    assert (found, ins) == (c["found"], c["insns"])
    assert n == len(ins)
    if found:
        end = pca.search_state.after(t, c["path"] + ins)
        assert end is not None and pca.search_state.leaf_check(t, end)


def test_completion_matches_oracle_on_generated_tasks(pca, oracle):
    """Seeded generated tasks (1-2 inputs, 8-32 tests) and random start
    states 0-2 instructions deep with a budget of 1-3, against the oracle's
    exact-key restatement; oracle calls above a few seconds are skipped."""
    from pca_amd import taskgen
    rng = random.Random(11)
    checked = found_n = 0
    for i in range(24):
        t = taskgen.chain_task(31, i, 3)
        ot = oracle.task_from_bytes(bytes(t))
        with pca.Solver(t) as sv:
            for _ in range(3):
                depth = rng.randrange(0, 3)
                path = []
                for _ in range(depth):
                    path.append(rng.choice(pca.search_state.gen_candidates(pca.search_state.after(t, path))))
                budget = rng.randrange(1, 4 if depth == 0 else 3)
                st = pca.search_state.after(t, path)
                ost = oracle.state_after(ot, path)
                print(f"[generated] task {i} path {path} max {depth + budget}", flush=True)
                t0 = time.perf_counter()
                want = oracle.complete(ot, ost, depth + budget, oracle.OEP_EXACT)
                if time.perf_counter() - t0 > 5.0:
                    continue
                found, n, ins = sv.exhaustive_complete(st, depth + budget)
                assert (found, ins) == (want[0], want[1]), (i, path)
                checked += 1
                found_n += found
    assert checked >= 50 and found_n >= 5


def test_completion_edge_cases(pca, solvers):
    c = next(x for x in CASES if x["kind"] == "root" and x["optimal_depth"] == 3)
    t, sv = solvers[c["task"]]
    root = pca.search_state.init_root(t)
    # budget 0 or negative: nothing to add
    assert sv.exhaustive_complete(root, 0) == (False, 0, [])
    deep = pca.search_state.after(t, c["insns"][:1])
    assert sv.exhaustive_complete(deep, 1) == (False, 0, [])
    # a solving start state needs no instruction, whatever the budget
    done = pca.search_state.after(t, c["insns"])
    assert sv.exhaustive_complete(done, 0) == (True, 0, [])
    # more than 8 instructions of budget is refused (the reference clamps at 16)
    with pytest.raises(pca.PcaGpuError):
        sv.exhaustive_complete(root, 9)
    # one handle, calls in a row: no state leaks between them
    # (python/test_exhaustive.py:98-108)
    first = sv.exhaustive_complete(root, 3)
    sv.exhaustive_complete(deep, 3)
    assert sv.exhaustive_complete(root, 3) == first == (True, 3, c["insns"])
