"""The allocation-free entry point (pcagpu_raw_hits_noalloc) against the
handle path it mirrors (begin_depth + submit + flush + pcagpu_raw_hits):
the same raw rows, bit for bit, on two calls in a row -- the second with a
smaller input, so stale hits or counters of the first would show."""
import random

import pytest

from conftest import golden_task_path

pytestmark = pytest.mark.gpu


def _handle_rows(pca, task, D, s, prefixes):
    with pca.Solver(task) as sv:
        return sv.raw_depth(D, s, prefixes)


def _planted_is_power2_32(pca, oracle):
    """7-instruction prefixes of the 32-test is_power2 task at depth 8 whose
    last three instructions begin proven 4-instruction solutions (fillers on
    unused registers first), plus random DFS paths: k = 1 rows with hits."""
    import json
    import os
    from conftest import GOLDEN
    t = pca.Task.load(golden_task_path("is_power2_32"))
    with open(os.path.join(GOLDEN, "exact", "is_power2_32_d4.json")) as f:
        sols = json.load(f)["solutions"]
    rng = random.Random(23)
    out = []
    for prog in sols[::6]:
        k = prog[1:5]
        used = set()
        for c in k:
            used.add((c >> 8) & 7)
            if ((c >> 11) & 31) not in (16, 18, 19):
                used.update(((c >> 5) & 7, (c >> 2) & 7))
        free = [r for r in range(1, 8) if r not in used]
        ins = [(16 << 11) | (free[q % len(free)] << 8) | rng.choice([0, 1, 255]) for q in range(4)] + k[:3]
        live = 1 << 1
        for c in ins:
            if ((c >> 11) & 31) not in (13, 19):
                live |= 1 << ((c >> 8) & 7)
        out.append((tuple(ins), live & 0xFE))
    for _ in range(40):
        live, ins = 1 << 1, []
        for pos in range(7):
            c = rng.choice(pca.gen_candidates(live, pos, 8))
            ins.append(c)
            if ((c >> 11) & 31) not in (13, 19):
                live |= 1 << ((c >> 8) & 7)
        out.append((tuple(ins), live & 0xFE))
    rng.shuffle(out)
    return t, out


@pytest.mark.parametrize("case", ["is_power2_32_d8_s7", "abs_d3_s2", "gen_chain_016_d3_s1", "ident0_d2_s1"])
def test_noalloc_equals_handle_path_twice(pca, oracle, case):
    if case == "is_power2_32_d8_s7":
        t, pre = _planted_is_power2_32(pca, oracle)
        D, s = 8, 7
    else:
        name, d, sp = case.rsplit("_", 2)
        D, s = int(d[1:]), int(sp[1:])
        t = pca.Task.load(golden_task_path(name))
        ot = oracle.task_from_bytes(bytes(t))
        pre = oracle.frontier(ot, D, s)
        pre = random.Random(31).sample(pre, min(len(pre), 12000))
    # both calls hold prefixes with hits: the owners of the handle path's
    # rows first, the rest after
    owners = sorted({r[0] for r in _handle_rows(pca, t, D, s, pre)})
    rest = [p for i, p in enumerate(pre) if i not in set(owners)]
    hit_pre = [pre[i] for i in owners]
    big = hit_pre + rest
    small = hit_pre[:max(1, len(hit_pre) // 2)] + rest[:len(rest) // 4]
    print(f"[noalloc] {case} D={D} s={s} prefixes {len(big)} then {len(small)}, {len(hit_pre)} with hits",
          flush=True)
    want_big = _handle_rows(pca, t, D, s, big)
    want_small = _handle_rows(pca, t, D, s, small)
    got_big = pca.raw_hits_noalloc(t, D, s, big)
    got_small = pca.raw_hits_noalloc(t, D, s, small)
    assert got_big == want_big
    assert got_small == want_small
    assert len(big) > len(small)
    assert len(want_big) >= len(want_small) > 0
    assert len(want_big) > len(want_small) or len(hit_pre) == 1


def test_noalloc_rejects_sizes_beyond_its_maximum(pca):
    t = pca.Task.load(golden_task_path("abs"))
    too_many = [((0, 0), 0b10)] * 16385
    with pytest.raises(pca.PcaGpuError) as e:
        pca.raw_hits_noalloc(t, 3, 2, too_many)
    assert e.value.code == -5
    # k = 3 below more prefixes than the block array holds (about 2,500
    # blocks below each root)
    prefixes = [((), 0b10)] * 200
    with pytest.raises(pca.PcaGpuError) as e:
        pca.raw_hits_noalloc(t, 3, 0, prefixes)
    assert e.value.code == -5
