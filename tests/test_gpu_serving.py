"""Serving: continuous batching over one persistent multi-task scan."""
import pytest

from conftest import golden_task_path

pytestmark = pytest.mark.gpu

GOLDEN = [("negate", 3), ("double", 3), ("add", 3), ("abs", 3), ("relu", 3), ("sign", 2), ("clamp", 2),
          ("is_power2", 2), ("ident0", 3), ("negate_16", 3), ("max", 2)]


def _tasks(pca):
    from pca_amd import taskgen
    tasks = [(pca.Task.load(golden_task_path(n)), d) for n, d in GOLDEN]
    tasks += [(taskgen.shallow_task(3, i), 4 + i % 5) for i in range(16)]
    return tasks


def _solo(pca, task, budget):
    with pca.Solver(task) as s:
        r = s.solve(budget, cap=4096)
    return r.optimal_depth, r.num_solutions, [x[0] for x in r.solutions]


def oracle_exact(oracle, task, budget):
    """The oracle's answer to a request: the first depth <= budget with
    programs, solved one depth at a time (exact-key frontier, last position
    on the VM), its program count and the first 4096 programs."""
    ot = oracle.task_from_bytes(bytes(task))
    for d in range(budget + 1):
        sols, _, _ = oracle.solve_at_depth(ot, d)
        if sols:
            return d, len(sols), sols[:4096]
    return -1, 0, []


@pytest.fixture(scope="module")
def solo(pca, oracle):
    """Every request's expected answer from the oracle; the solo HIP solve
    must give it too."""
    tasks = _tasks(pca)
    want = [oracle_exact(oracle, t, d) for t, d in tasks]
    for (t, d), w in zip(tasks, want):
        assert _solo(pca, t, d) == w
    return tasks, want


@pytest.mark.parametrize("kw", [{}, {"warm_slots": 0, "prefixes_per_item": 2048, "step_prefixes": 4096},
                                {"warm_slots": 8, "step_prefixes": 1}],
                         ids=["default", "small-shares", "warm-no-step-share"])
def test_server_mix_equals_solo(pca, solo, kw):
    """Every request's answer equals the oracle's (and pcagpu_solve's on the
    task alone), with the default shares, with small shares (many steps per depth, the test
    order adapting across them) and with the step share switched off."""
    from pca_amd.serving import Server, DONE
    tasks, want = solo
    print(f"[serving] {len(tasks)} tasks {kw}", flush=True)
    with Server(max_slots=8, **kw) as sv:
        ids = [sv.submit(t, d) for t, d in tasks[:10]]
        for _ in range(3):
            sv.step()                      # the rest arrive while these run
        ids += [sv.submit(t, d) for t, d in tasks[10:]]
        recs = {r["id"]: r for r in sv.drain()}
        st = sv.stats()
        for i, rid in enumerate(ids):
            r = recs[rid]
            assert r["status"] == DONE
            sols, n = sv.solutions(rid)
            got = (r["optimal_depth"], n, [x[0] for x in sols])
            assert got == want[i], (i, got[:2], want[i][:2])
    assert st["items"] > st["launches"], "tasks did not share launches"
    assert st["slots"] <= 8
    # host time by phase is accounted
    assert st["stage_ns"] > 0 and st["wait_ns"] > 0 and st["resolve_ns"] > 0 and st["admit_ns"] > 0


def test_server_hit_overflow_rescan_matches_oracle(pca, oracle):
    """A slot whose hits overflow its device buffer rescans that batch alone
    with a larger one: with 64 records per slot, max (243 raw hits at depth
    3) still gets the oracle's set, next to requests that do not overflow."""
    from pca_amd.serving import Server, DONE
    reqs = [(pca.Task.load(golden_task_path("add")), 2), (pca.Task.load(golden_task_path("abs")), 3),
            (pca.Task.load(golden_task_path("negate")), 2), (pca.Task.load(golden_task_path("max")), 3)]
    want = [oracle_exact(oracle, t, d) for t, d in reqs]
    print(f"[serving-overflow] {[w[:2] for w in want]}", flush=True)
    with Server(max_slots=4, hit_capacity=64) as sv:
        ids = [sv.submit(t, d) for t, d in reqs]
        recs = {r["id"]: r for r in sv.drain()}
        for rid, w in zip(ids, want):
            assert recs[rid]["status"] == DONE
            sols, n = sv.solutions(rid)
            assert (recs[rid]["optimal_depth"], n, [x[0] for x in sols]) == w
        assert sv.stats()["rescans"] >= 1


def test_server_failed_step_fails_its_requests_and_recovers(pca, oracle):
    """A step that fails after staging (forced from the host: fail_step)
    fails exactly the requests staged in it; the others finish with the
    oracle's answers, and the same tasks submitted again are answered
    correctly by the re-tasked slots."""
    from pca_amd.serving import Server, DONE, FAILED
    from pca_amd import PcaGpuError
    reqs = [(pca.Task.load(golden_task_path(n)), d) for n, d in
            (("abs", 3), ("relu", 3), ("negate", 3), ("add", 3), ("double", 3), ("max", 2))]
    want = [oracle_exact(oracle, t, d) for t, d in reqs]
    print("[serving-fail-step]", flush=True)
    with Server(max_slots=3, warm_slots=3, fail_step=2, prefixes_per_item=4096, step_prefixes=1) as sv:
        ids = [sv.submit(t, d) for t, d in reqs]
        recs, errors = {}, 0
        for _ in range(10000):
            try:
                pending = sv.step()
            except PcaGpuError:
                errors += 1
                continue
            for r in sv.poll():
                recs[r["id"]] = r
            if pending == 0:
                break
        for r in sv.poll():
            recs[r["id"]] = r
        assert errors == 1
        failed = [i for i, rid in enumerate(ids) if recs[rid]["status"] == FAILED]
        assert failed
        for i, rid in enumerate(ids):
            if i in failed:
                continue
            assert recs[rid]["status"] == DONE
            sols, n = sv.solutions(rid)
            assert (recs[rid]["optimal_depth"], n, [x[0] for x in sols]) == want[i]
        again = {sv.submit(*reqs[i]): i for i in failed}
        for r in sv.drain():
            recs[r["id"]] = r
        for rid, i in again.items():
            assert recs[rid]["status"] == DONE
            sols, n = sv.solutions(rid)
            assert (recs[rid]["optimal_depth"], n, [x[0] for x in sols]) == want[i]


@pytest.mark.parametrize("warm,max_slots,want", [(None, 64, 4), (0, 64, 0), (8, 3, 3), (2, 64, 2)])
def test_server_warm_slots(pca, warm, max_slots, want):
    """Warm slots are opened by server_open (none when switched off, at most
    max_slots); the first request is re-tasked into one instead of opening
    a slot, and its answer is the solver's."""
    from pca_amd.serving import Server, DONE
    t, d = _tasks(pca)[0]
    kw = {} if warm is None else {"warm_slots": warm}
    with Server(max_slots=max_slots, **kw) as sv:
        assert sv.stats()["slots_opened"] == want
        rid = sv.submit(t, d)
        recs = {r["id"]: r for r in sv.drain()}
        st = sv.stats()
        sols, n = sv.solutions(rid)
    assert recs[rid]["status"] == DONE
    assert (recs[rid]["optimal_depth"], n, [x[0] for x in sols]) == _solo(pca, t, d)
    assert st["slots_opened"] == max(want, 1)


@pytest.mark.parametrize("max_slots,budget", [(2, 200 << 30), (8, 1)])
def test_server_admission_limits(pca, max_slots, budget):
    from pca_amd.serving import Server, DONE
    tasks = _tasks(pca)[:12]
    with Server(max_slots=max_slots, hbm_budget=budget) as sv:
        ids = [sv.submit(t, d) for t, d in tasks]
        recs = {r["id"]: r for r in sv.drain()}
        st = sv.stats()
    assert all(recs[i]["status"] == DONE for i in ids)
    assert st["slots"] <= (max_slots if budget > 1 else 1)


def test_server_rejects_bad_task_without_stalling(pca):
    from pca_amd.serving import Server, DONE, FAILED
    bad = pca.Task.load(golden_task_path("negate"))
    bad.num_outputs = 2                     # multi-output: unsupported
    good = pca.Task.load(golden_task_path("double"))
    with Server(max_slots=2) as sv:
        b = sv.submit(bad, 2)
        g = sv.submit(good, 2)
        recs = {r["id"]: r for r in sv.drain()}
    assert recs[b]["status"] == FAILED and recs[g]["status"] == DONE


def test_suite_budget4_server_equals_one_at_a_time(pca, oracle):
    """Config-3 generator at a depth budget of 4: the tasks that need depth 4
    give frontiers whose scans fill the chip alone (their own launches,
    capped automatic split) next to many small shared items; every task's
    optimum, solution count and programs equal the one-at-a-time solver's,
    and the oracle's for every task whose optimum is at most 3."""
    from pca_amd import taskgen, serving
    tasks = [taskgen.random_task(11, i) for i in range(150)]
    a, wa, st = serving.suite_run(tasks, 4, max_slots=256, prefixes_per_item=1 << 20)
    b, wb = serving.sequential_suite(tasks, 4)
    print(f"[suite] solved {sum(x[0] >= 0 for x in a)}/150 at budget 4, depth 4: {sum(x[0] == 4 for x in a)}; "
          f"server {wa:.2f}s in {st['launches']} launches, one at a time {wb:.2f}s", flush=True)
    assert a == b
    assert sum(x[0] == 4 for x in a) >= 1
    checked = 0
    for t, got in zip(tasks, a):
        if 0 <= got[0] <= 3:
            assert got == oracle_exact(oracle, t, 3)
            checked += 1
    assert checked >= 100
