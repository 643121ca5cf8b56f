"""The benchmark's workload at its full size -- is_power2_32 (32 tests),
kernel depth 8, split 5, 3-instruction suffixes -- where the oracle cannot
follow: properties that hold whatever the size."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def window(pca):
    from pca_amd import taskgen
    t = taskgen.is_power2_32(42)
    prefixes, _, _ = pca.host_frontier(t, 8, 5, cap=8)
    assert len(prefixes) == 8
    return t, prefixes


def _scan(pca, t, prefixes, **kw):
    with pca.Solver(t, **kw) as sv:
        hits = sv.raw_depth(8, 5, prefixes)
        st = sv.stats()
    return hits, {k: st[k] for k in ("candidates", "suffixes", "parents", "prefixes")}


def test_full_size_window_is_schedule_independent(pca, window):
    """Eight DFS-ordered depth-5 prefixes scanned to depth 8 with every tile
    the autotuner may pick, with the adaptive test order on and off, and
    through the graph and the eager flush: the same candidate, suffix and
    parent counts and the same raw hits (the schedule decides only how the
    work is cut, never what is found)."""
    t, prefixes = window
    ref = None
    for tile in [(64, 3), (32, 2), (16, 1)]:
        for fixed in (False, True):
            for eager in (False, True):
                print(f"[full-size] tile={tile} fixed_order={fixed} eager={eager}", flush=True)
                got = _scan(pca, t, prefixes, scan_tile=tile, fixed_test_order=fixed, eager_flush=eager)
                assert got[1]["candidates"] > 10 ** 9
                if ref is None:
                    ref = got
                assert got == ref, (tile, fixed, eager)


def test_full_size_window_splits_into_prefix_batches(pca, window):
    """The same eight prefixes scanned one at a time: counts add up and the
    hits are the same rows (prefix index shifted by the batch offset)."""
    t, prefixes = window
    hits, counts = _scan(pca, t, prefixes)
    total = {k: 0 for k in counts}
    rows = []
    for i, p in enumerate(prefixes):
        h, c = _scan(pca, t, [p])
        rows += [(i, f, m, j) for (_, f, m, j) in h]
        for k in total:
            total[k] += c[k]
    assert total == counts
    assert sorted(rows) == sorted(hits)
