"""The reference's branchless search state (src/search_state.h:19-25) and
the host calls its learned-search loop makes on it (python/search_eval.py:
85-187, PCALib): root, step, leaf check, candidate list.  These are host
helpers of the C ABI (include/pcagpu.h) for building and checking states;
the exhaustive completion itself runs on the GPU
(Solver.exhaustive_complete, pcagpu_complete).

    from pca_amd import Task, Solver, search_state as ss
    task = Task.load("task.json")
    st = ss.step(task, ss.init_root(task), insn)
    with Solver(task) as sv:
        found, n, insns = sv.exhaustive_complete(st, max_depth=4)
"""
import ctypes as C

from ._lib import PcaGpuError, SearchState, Task, lib

__all__ = ["SearchState", "init_root", "step", "leaf_check", "gen_candidates", "after"]


def init_root(task):
    """search_state_init_root (src/search_state.c:34-46)."""
    st = SearchState()
    lib().pcagpu_state_init_root(C.byref(task), C.byref(st))
    return st


def step(task, state, insn):
    """search_state_step (src/search_state.c:49-187): the child, or None if
    the instruction is rejected (not branchless, or a DIV/MOD by zero)."""
    out = SearchState()
    if not lib().pcagpu_state_step(C.byref(task), C.byref(state), insn, C.byref(out)):
        return None
    return out


def leaf_check(task, state):
    """search_state_leaf_outputs (src/search_state.c:223-239): mask of live
    registers equal to the target on every test."""
    return int(lib().pcagpu_state_leaf_outputs(C.byref(task), C.byref(state)))


def gen_candidates(state_or_live):
    """search_gen_candidates_branchless (src/search_state.c:280-330)."""
    live = state_or_live.live if isinstance(state_or_live, SearchState) else int(state_or_live)
    out = (C.c_uint16 * 4096)()
    n = lib().pcagpu_gen_candidates_branchless(out, 4096, live)
    if n < 0:
        raise PcaGpuError(n, "gen_candidates_branchless")
    return list(out[:n])


def after(task, insns):
    """The root state stepped through insns (None if a step is rejected)."""
    if not isinstance(task, Task):
        raise TypeError("task must be a pca_amd.Task")
    st = init_root(task)
    for c in insns:
        st = step(task, st, c)
        if st is None:
            return None
    return st
