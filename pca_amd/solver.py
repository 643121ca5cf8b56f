"""Python host over the C ABI: the reference's hybrid solver loop
(gpu/hybrid.cu:631-836) with the device doing every candidate evaluation.

``Solver`` owns one device handle (one resident task).  ``solve`` runs the
IDDFS inside the library; ``hybrid_depth`` drives the drop-in seam
(begin/submit/flush/end) from Python the way hybrid.cu's main() drives its
kernels, which is what the parity tests exercise.
"""
import ctypes as C
from dataclasses import dataclass, field

from ._lib import (PcaGpuError, Task, Prefix, Solution, Stats, Opts, SolveOpts, SolveRes, lib,
                   MAX_KDEPTH)


@dataclass
class SolveResult:
    optimal_depth: int
    solutions: list               # list of (prog list, out_reg, prefix_rank, cand path)
    num_solutions: int
    depths_done: int
    timed_out: bool
    seconds: float
    candidates: list = field(default_factory=list)
    prefixes: list = field(default_factory=list)
    depth_seconds: list = field(default_factory=list)


def _sol_tuple(s):
    return (s.program(), int(s.out_reg), int(s.prefix_rank), tuple(s.cand[:s.kernel_len]),
            bool(s.cpu_verified))


class Solver:
    def __init__(self, task, device=0, batch_capacity=0, hit_capacity=0, waves_per_cu=0, fixed_test_order=False,
                 scan_tile=None, eager_flush=False, test_flags=0, per_child=False):
        """scan_tile: None (autotuned) or (block_kids 1..64, mask_tests 1..6).
        per_child: scan every block child's last position on its own instead
        of sharing the value groups that do not read its register (the same
        results; the slower form, kept to check the shared one).
        test_flags: testing only (see pcagpu_opts_t.test_flags)."""
        if not isinstance(task, Task):
            raise TypeError("task must be a pca_amd.Task")
        self.task = task
        opts = Opts(device=device, batch_capacity=batch_capacity, hit_capacity=hit_capacity,
                    waves_per_cu=waves_per_cu, fixed_test_order=1 if fixed_test_order else 0,
                    scan_tile=(0 if scan_tile is None else (scan_tile[0] | (scan_tile[1] << 8))) |
                    ((1 << 16) if per_child else 0),
                    eager_flush=1 if eager_flush else 0, test_flags=test_flags)
        h = C.c_void_p()
        rc = lib().pcagpu_open(C.byref(task), C.byref(opts), C.byref(h))
        if rc < 0:
            raise PcaGpuError(rc, "pcagpu_open")
        self._h = h

    # context manager ---------------------------------------------------
    def close(self):
        if self._h:
            lib().pcagpu_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def device_name(self):
        return lib().pcagpu_device_name(self._h).decode()

    def stats(self):
        st = Stats()
        rc = lib().pcagpu_get_stats(self._h, C.byref(st))
        if rc < 0:
            raise PcaGpuError(rc, "get_stats")
        return st.as_dict()

    def reset_stats(self):
        lib().pcagpu_reset_stats(self._h)

    def retask(self, task):
        """Points this resident handle at another task (pcagpu_retask): no
        device allocation, only the task record is uploaded."""
        if not isinstance(task, Task):
            raise TypeError("task must be a pca_amd.Task")
        rc = lib().pcagpu_retask(self._h, C.byref(task))
        if rc < 0:
            raise PcaGpuError(rc, "pcagpu_retask")
        self.task = task

    # whole solve -------------------------------------------------------
    def solve(self, max_depth, start_depth=0, split=-1, producer="gpu", time_limit=0.0,
              cap=4096, verbose=False):
        o = SolveOpts(max_depth=max_depth, start_depth=start_depth, split=split,
                      producer=0 if producer == "gpu" else 1, verbose=int(verbose),
                      time_limit=time_limit)
        res = SolveRes()
        out = (Solution * cap)()
        rc = lib().pcagpu_solve(self._h, C.byref(o), C.byref(res), out, cap)
        if rc < 0:
            raise PcaGpuError(rc, "pcagpu_solve")
        n = min(res.num_solutions, cap)
        return SolveResult(optimal_depth=res.optimal_depth,
                           solutions=[_sol_tuple(out[i]) for i in range(n)],
                           num_solutions=int(res.num_solutions), depths_done=res.depths_done,
                           timed_out=bool(res.timed_out), seconds=res.seconds,
                           candidates=list(res.candidates), prefixes=list(res.prefixes),
                           depth_seconds=list(res.depth_seconds))

    # exhaustive leaf completion (src/search_state.c:418-450) -------------
    def exhaustive_complete(self, state, max_depth):
        """search_exhaustive_complete on the GPU, as PCALib.exhaustive_complete
        (python/search_eval.py:171-181): (found, sol_len, [insns])."""
        ins = (C.c_uint16 * MAX_KDEPTH)()
        n = C.c_int()
        rc = lib().pcagpu_complete(self._h, C.byref(state), max_depth, ins, C.byref(n))
        if rc < 0:
            raise PcaGpuError(rc, "pcagpu_complete")
        if rc == 0:
            return False, 0, []
        return True, n.value, list(ins[:n.value])

    def exhaustive_states_explored(self):
        """Children generated by the last exhaustive_complete."""
        return int(lib().pcagpu_complete_states_explored(self._h))

    # the drop-in seam ----------------------------------------------------
    def frontier(self, target_depth, split, cap=1 << 24):
        """GPU wavefront producer: exact-OEP survivors at the split, DFS order."""
        n = lib().pcagpu_frontier(self._h, target_depth, split, None, 0)
        if n < 0:
            raise PcaGpuError(n, "pcagpu_frontier")
        buf = (Prefix * max(n, 1))()
        m = lib().pcagpu_frontier(self._h, target_depth, split, buf, n)
        if m < 0:
            raise PcaGpuError(m, "pcagpu_frontier")
        return [(tuple(buf[i].insns[:split]), buf[i].live) for i in range(n)]

    def raw_depth(self, target_depth, split, prefixes, first_rank=0):
        """Runs one depth's scan and returns the raw device hits (before exact
        OEP) as (prefix index, flags, mask, j tuple) rows in canonical order;
        the depth is then closed."""
        rc = lib().pcagpu_begin_depth(self._h, target_depth, split)
        if rc < 0:
            raise PcaGpuError(rc, "begin_depth")
        arr = (Prefix * max(len(prefixes), 1))()
        for i, (ins, live) in enumerate(prefixes):
            for q, v in enumerate(ins):
                arr[i].insns[q] = v
            arr[i].num_insns = split
            arr[i].live = live
        if prefixes:
            rc = lib().pcagpu_submit(self._h, arr, len(prefixes), first_rank)
            if rc < 0:
                raise PcaGpuError(rc, "submit")
        rc = lib().pcagpu_flush(self._h)
        if rc < 0:
            raise PcaGpuError(rc, "flush")
        n = lib().pcagpu_raw_hits(self._h, None, 0)
        rows = (C.c_uint16 * (14 * max(n, 1)))()
        lib().pcagpu_raw_hits(self._h, rows, n)
        k = target_depth - split
        out = []
        for i in range(n):
            r = rows[i * 14:(i + 1) * 14]
            pidx = r[0] | (r[1] << 16) | (r[2] << 32) | (r[3] << 48)
            out.append((pidx, r[4], r[5], tuple(r[6:6 + k])))
        lib().pcagpu_end_depth(self._h, None, 0)
        return out

    def hybrid_depth(self, target_depth, split, prefixes, first_rank=0, cap=1 << 16):
        """One IDDFS depth through begin/submit/flush/end with caller-made
        prefixes [(insns tuple, live)] in DFS order."""
        rc = lib().pcagpu_begin_depth(self._h, target_depth, split)
        if rc < 0:
            raise PcaGpuError(rc, "begin_depth")
        arr = (Prefix * max(len(prefixes), 1))()
        for i, (ins, live) in enumerate(prefixes):
            for q, v in enumerate(ins):
                arr[i].insns[q] = v
            arr[i].num_insns = split
            arr[i].live = live
        if prefixes:
            rc = lib().pcagpu_submit(self._h, arr, len(prefixes), first_rank)
            if rc < 0:
                raise PcaGpuError(rc, "submit")
        rc = lib().pcagpu_flush(self._h)
        if rc < 0:
            raise PcaGpuError(rc, "flush")
        out = (Solution * cap)()
        n = lib().pcagpu_end_depth(self._h, out, cap)
        if n < 0:
            raise PcaGpuError(n, "end_depth")
        return [_sol_tuple(out[i]) for i in range(min(n, cap))], int(n)


def raw_hits_noalloc(task, target_depth, split, prefixes, cap=None):
    """pcagpu_raw_hits_noalloc: the raw device hits below `prefixes`
    [(insns, live)] at (target_depth, split), computed without a handle and
    without allocating device memory, as (prefix index, flags, mask, j tuple)
    rows in canonical order -- the rows Solver.raw_depth returns."""
    n = len(prefixes)
    arr = (Prefix * max(n, 1))()
    for i, (ins, live) in enumerate(prefixes):
        for q, v in enumerate(ins):
            arr[i].insns[q] = v
        arr[i].num_insns = split
        arr[i].live = live
    if cap is None:
        cap = 1 << 18
    rows = (C.c_uint16 * (14 * max(cap, 1)))()
    m = lib().pcagpu_raw_hits_noalloc(C.byref(task), target_depth, split, arr, n, rows, cap)
    if m < 0:
        raise PcaGpuError(m, "pcagpu_raw_hits_noalloc")
    k = target_depth - split
    out = []
    for i in range(min(m, cap)):
        r = rows[i * 14:(i + 1) * 14]
        out.append((r[0] | (r[1] << 16), r[4], r[5], tuple(r[6:6 + k])))
    return out

