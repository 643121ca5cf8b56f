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


@pytest.fixture(scope="module")
def solo(pca):
    tasks = _tasks(pca)
    return tasks, [_solo(pca, t, d) for t, d in tasks]


@pytest.mark.parametrize("kw", [{}, {"warm_slots": 0, "prefixes_per_item": 2048, "step_prefixes": 4096},
                                {"warm_slots": 8, "step_prefixes": 1}],
                         ids=["default", "small-shares", "warm-no-step-share"])
def test_server_mix_equals_solo(pca, solo, kw):
    """Every request's answer equals pcagpu_solve's on the task alone, with
    the default shares, with small shares (many steps per depth, the test
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


def test_suite_budget4_server_equals_one_at_a_time(pca):
    """Config-3 generator at a depth budget of 4: the tasks that need depth 4
    give frontiers whose scans fill the chip alone (their own launches,
    capped automatic split) next to many small shared items; every task's
    optimum and solution count equal the one-at-a-time solver's."""
    from pca_amd import taskgen, serving
    tasks = [taskgen.random_task(11, i) for i in range(150)]
    a, wa, st = serving.suite_run(tasks, 4, max_slots=256, prefixes_per_item=1 << 20)
    b, wb = serving.sequential_suite(tasks, 4)
    print(f"[suite] solved {sum(x[0] >= 0 for x in a)}/150 at budget 4, depth 4: {sum(x[0] == 4 for x in a)}; "
          f"server {wa:.2f}s in {st['launches']} launches, one at a time {wb:.2f}s", flush=True)
    assert a == b
    assert sum(x[0] == 4 for x in a) >= 1
