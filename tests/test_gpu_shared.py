"""The last position shared per block (the default scan tile) against every
child scanned on its own (per_child=True, pinned to the oracle by the
parity tests): the same raw hits, rows and flags, and the same candidate,
suffix and parent counts, on suffixes of 2 and 3 instructions -- planted
prefixes that reach solutions, random DFS paths, and generated tasks whose
tests hold zeros (zero divisors, where a DIV / MOD group's rd = x value
reads the child's register)."""
import random

import pytest

from conftest import golden_task_path
from test_gpu_deep import _fixture, _kernel_of, _planted_prefixes, _random_prefixes

pytestmark = pytest.mark.gpu

COUNTS = ("candidates", "suffixes", "parents", "prefixes")


def _both(pca, t, D, s, prefixes):
    out = []
    for per_child in (True, False):
        with pca.Solver(t, per_child=per_child, fixed_test_order=True) as sv:
            raw = sv.raw_depth(D, s, prefixes)
            st = sv.stats()
        out.append((raw, {k: st[k] for k in COUNTS}, st["tests_evaluated"], st["scan_tile"] >> 16))
    (r0, c0, e0, m0), (r1, c1, e1, m1) = out
    assert (m0, m1) == (1, 0), (m0, m1)
    return r0, c0, e0, r1, c1, e1


@pytest.mark.parametrize("s", [6, 5])
def test_shared_equals_per_child_on_planted_solutions(pca, s):
    """is_power2_32 at kernel depth 8 below prefixes that end in the head of
    proven 4-instruction solutions (fillers first): k = 2 and k = 3 suffixes
    that contain solutions."""
    t = pca.Task.load(golden_task_path("is_power2_32"))
    fx = _fixture("is_power2_32", 4)
    prefixes = []
    for k in [_kernel_of(p, t.num_inputs, 4) for p in fx["solutions"][::6]]:
        body = k[:s - 4]
        prefixes += _planted_prefixes(t, k, body, s - len(body), 2)
    print(f"[shared] planted is_power2_32 D=8 s={s} prefixes={len(prefixes)}", flush=True)
    r0, c0, e0, r1, c1, e1 = _both(pca, t, 8, s, prefixes)
    assert r1 == r0
    assert c1 == c0
    assert any(not (f & 1) for _, f, _, _ in r0), "no passing hit in the sample"
    print(f"[shared] hits={len(r0)} tests per-child={e0} shared={e1}", flush=True)


@pytest.mark.parametrize("seed", range(6))
def test_shared_equals_per_child_random_paths(pca, seed):
    """Random DFS paths of generated tasks (inputs often zero), k = 2 and
    k = 3 at kernel depths 4-6: every raw row identical."""
    from pca_amd import taskgen
    t = taskgen.random_task(100 + seed, seed)
    rng = random.Random(seed)
    for i in range(t.num_tests):
        for q in range(t.num_inputs):
            if rng.random() < 0.4:
                t.tests[i].inp[q] = rng.choice([0, 0, 1, 2, 0xFFFF])
        # targets: one input or zero, so some value groups survive many tests
        t.tests[i].out[0] = t.tests[i].inp[0] if rng.random() < 0.7 else 0
    D = 4 + seed % 3
    for s in (D - 2, D - 3):
        prefixes = _random_prefixes(pca, t, D, s, 40, seed=seed * 7 + s)
        print(f"[shared] random_task seed={seed} D={D} s={s} tests={t.num_tests}", flush=True)
        r0, c0, e0, r1, c1, e1 = _both(pca, t, D, s, prefixes)
        assert r1 == r0, (seed, D, s)
        assert c1 == c0


def test_shared_equals_per_child_full_size_window(pca):
    """The bench's workload shape (is_power2_32, depth 8, split 5) on eight
    DFS-ordered prefixes: identical counts and rows."""
    from pca_amd import taskgen
    t = taskgen.is_power2_32(42)
    prefixes, _, _ = pca.host_frontier(t, 8, 5, cap=8)
    r0, c0, e0, r1, c1, e1 = _both(pca, t, 8, 5, prefixes)
    assert r1 == r0
    assert c1 == c0
    assert c0["candidates"] > 10 ** 9
    print(f"[shared] full-size tests per-child={e0} shared={e1} ({e1 / e0:.4f})", flush=True)
