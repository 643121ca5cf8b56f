"""Failure and limit paths of the device search (forced from the host; every
kernel still exits normally): a hit buffer far too small (regrown and the
launch redone, never a dropped hit), a cycle limit that binds (programs
that run out of cycles cannot pass, on some tests only when a branch is not
taken), and the wavefront's refusals -- a table that overflows and 64-bit
keys that collide return errors instead of a wrong frontier."""
import json
import os

import pytest

from conftest import GOLDEN, golden_task_path

pytestmark = pytest.mark.gpu


def _fixture(name, d):
    with open(os.path.join(GOLDEN, "exact", f"{name}_d{d}.json")) as f:
        return json.load(f)


def _oracle_exact(oracle, task, budget):
    ot = oracle.task_from_bytes(bytes(task))
    for d in range(budget + 1):
        sols, _, _ = oracle.solve_at_depth(ot, d)
        if sols:
            return d, len(sols), sols
    return -1, 0, []


@pytest.mark.parametrize("name,d,cap", [("max", 3, 64), ("gen_chain_050", 3, 16), ("is_power2_32", 4, 1)])
def test_tiny_hit_buffer_regrows_and_keeps_every_hit(pca, name, d, cap):
    """hit_capacity of 1-64 records: each overflowing launch is repeated
    with a larger buffer; the solution set is the oracle's fixture and the
    raw hit count equals the default buffer's."""
    want = _fixture(name, d)
    t = pca.Task.load(golden_task_path(name))
    print(f"[hit-regrow] {name} d{d} cap={cap}", flush=True)
    with pca.Solver(t, hit_capacity=cap) as sv:
        res = sv.solve(d, cap=1 << 16)
        raw_small = sv.stats()["raw_hits"]
    with pca.Solver(t) as sv:
        sv.solve(d, cap=1 << 16)
        raw_default = sv.stats()["raw_hits"]
    assert res.optimal_depth == want["optimal_depth"]
    assert [s[0] for s in res.solutions] == want["solutions"]
    assert raw_small == raw_default > cap


@pytest.mark.parametrize("name,budget,cycles", [("abs", 3, 5), ("abs", 4, 6), ("relu", 3, 5), ("max", 3, 6),
                                                 ("max", 3, 7), ("add", 2, 4)])
def test_binding_cycle_limit_matches_oracle(pca, oracle, name, budget, cycles):
    """max_cycles at or below the optimal program's length: the device must
    drop exactly the programs the VM times out (its dead flag and the
    last-position cycle check), as the oracle does -- here every optimal
    program or none of them (one that skipped instructions on every test
    would have a shorter equivalent)."""
    t = pca.Task.load(golden_task_path(name))
    t.max_cycles = cycles
    want = _oracle_exact(oracle, t, budget)
    print(f"[cycles] {name} budget {budget} max_cycles {cycles}: oracle D*={want[0]} n={want[1]}", flush=True)
    for split in (-1, 0):
        with pca.Solver(t) as sv:
            res = sv.solve(budget, split=split, cap=1 << 16)
        assert (res.optimal_depth, res.num_solutions, [s[0] for s in res.solutions]) == want
        assert all(s[4] for s in res.solutions)


@pytest.mark.parametrize("flag,code", [(1, -5), (2, -6)])
def test_wavefront_refuses_overflow_and_collisions(pca, flag, code):
    """A 64-slot table (flag 1) cannot hold a level: PCAGPU_EOVERFLOW.  Keys
    cut to 3 bits (flag 2) make distinct states share keys; the full-state
    confirmation sees it: PCAGPU_ECOLLISION.  Never a silent wrong frontier."""
    t = pca.Task.load(golden_task_path("abs"))
    print(f"[wavefront-refusal] flag {flag}", flush=True)
    with pca.Solver(t, test_flags=flag) as sv:
        with pytest.raises(pca.PcaGpuError) as e:
            sv.frontier(3, 2)
        assert e.value.code == code
        with pytest.raises(pca.PcaGpuError) as e:
            sv.solve(3)
        assert e.value.code == code
    with pca.Solver(t) as sv:                        # the same level, normal table
        assert len(sv.frontier(3, 2)) == 29244
