"""Leaf completion on the CPU: the oracle's restatement of
search_exhaustive_complete (src/search_state.c:339-445) pinned against the
reference library's own answers (tests/golden/complete/cases.json, written
by tests/golden/make_complete_golden.py from oracle/_ref/libpca.so), the
manifest's optimal depths and an unpruned brute force; and the product's
host helpers for search states against the oracle.  No device calls."""
import ctypes as C
import json
import os
import random

import pytest

from conftest import GOLDEN

CDIR = os.path.join(GOLDEN, "complete")
with open(os.path.join(CDIR, "cases.json")) as _f:
    CASES = json.load(_f)["cases"]
# cases the oracle finishes in well under a second each
FAST = [c for c in CASES if c["seconds"] < 0.3]


def _task(oracle, name):
    return oracle.load_task(os.path.join(CDIR, "tasks", name + ".json"))


def _cid(c):
    return f"{c['task']}-{c['kind']}-{'.'.join(map(str, c['path']))}-{c['max_depth']}"


def test_fixture_set_covers_the_reference_test_plan():
    """python/test_exhaustive.py:53-110: root completion at the optimal depth,
    none below it, completion of the state after one instruction."""
    kinds = {c["kind"] for c in CASES}
    assert {"root", "below", "mid", "solved", "random", "heavy"} <= kinds
    assert sum(c["found"] for c in CASES) > 40 and sum(not c["found"] for c in CASES) > 40
    assert len(FAST) >= 100


@pytest.mark.parametrize("c", FAST, ids=_cid)
def test_oracle_ref32_equals_reference_library(oracle, c):
    """The oracle with the reference's own 32-bit OEP table returns the same
    instructions and the same explored-state count as the reference."""
    t = _task(oracle, c["task"])
    st = oracle.state_after(t, c["path"])
    found, ins, explored = oracle.complete(t, st, c["max_depth"], oracle.OEP_REF32)
    assert (found, ins, explored) == (c["found"], c["insns"], c["explored"])


@pytest.mark.parametrize("c", FAST, ids=_cid)
def test_oracle_exact_and_unpruned_agree_with_reference(oracle, c):
    """Exact OEP keys (the GPU's semantics) and no OEP at all give the
    reference's answer: the first completion in DFS order survives every
    prune by state on these cases."""
    t = _task(oracle, c["task"])
    st = oracle.state_after(t, c["path"])
    for mode in (oracle.OEP_EXACT, oracle.OEP_OFF):
        found, ins, _ = oracle.complete(t, st, c["max_depth"], mode)
        assert (found, ins) == (c["found"], c["insns"]), mode


def test_root_completion_has_the_manifest_depth(oracle):
    """data/manifest.json optimal depths 1..3 (the reference's test asserts
    sol_len == d, python/test_exhaustive.py:69-72): found at d with d
    instructions that solve when replayed, and not found within d-1."""
    seen = 0
    for c in CASES:
        if c["kind"] == "root" and c["optimal_depth"] <= 3 and c["seconds"] < 1.5:
            t = _task(oracle, c["task"])
            found, ins, _ = oracle.complete(t, oracle.state_root(t), c["optimal_depth"], oracle.OEP_EXACT)
            assert found and len(ins) == c["optimal_depth"]
            st = oracle.state_after(t, ins)
            assert st is not None and oracle.state_leaf(t, st)
            if c["optimal_depth"] >= 2:
                assert not oracle.complete(t, oracle.state_root(t), c["optimal_depth"] - 1, oracle.OEP_EXACT)[0]
            seen += 1
    assert seen >= 12


# ------------------------------------------------------ product host helpers
def test_search_state_layout(pca):
    assert C.sizeof(pca.search_state.SearchState) == 546   # search_state_t (src/search_state.h:19-25)


def test_branchless_candidates_match_oracle_all_masks(pca, oracle):
    for m in range(256):
        assert pca.search_state.gen_candidates(m) == oracle.gen_candidates_branchless(m)
    # the reference's list sizes (1 live input, 2, all live)
    assert [len(pca.search_state.gen_candidates(m)) for m in (0x02, 0x06, 0xFE)] == [214, 435, 2905]


def test_state_step_and_leaf_match_oracle(pca, oracle):
    """Random walks over the whole opcode space (DIV/MOD faults, non-branchless
    opcodes rejected) on the product's host step vs the oracle's."""
    rng = random.Random(5)
    names = sorted({c["task"] for c in CASES})
    for name in names:
        path = os.path.join(CDIR, "tasks", name + ".json")
        t = pca.Task.load(path)
        ot = oracle.load_task(path)
        a, b = pca.search_state.init_root(t), oracle.state_root(ot)
        assert bytes(a) == bytes(b)
        for _ in range(400):
            insn = rng.randrange(65536) if rng.random() < 0.5 else rng.choice(pca.search_state.gen_candidates(a))
            na, nb = pca.search_state.step(t, a, insn), oracle.state_step(ot, b, insn)
            assert (na is None) == (nb is None), hex(insn)
            if na is None:
                continue
            assert bytes(na) == bytes(nb)
            assert pca.search_state.leaf_check(t, na) == oracle.state_leaf(ot, nb)
            a, b = (na, nb) if na.depth < 6 else (pca.search_state.init_root(t), oracle.state_root(ot))
