"""ctypes wrapper of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, always as the checker or the baseline,
never as part of the product path.
"""
import ctypes as C
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
CLI_PATH = os.path.join(HERE, "oracle_enumerate")
REF_ENUM = os.path.join(HERE, "_ref", "enumerate")

OEP_OFF, OEP_EXACT, OEP_REF32 = 0, 1, 2


class OrTest(C.Structure):
    _fields_ = [("inp", C.c_uint16 * 8), ("out", C.c_uint16 * 8)]


class OrTask(C.Structure):
    _fields_ = [("name", C.c_char * 64), ("max_cycles", C.c_uint64), ("input_ports", C.c_uint8 * 8),
                ("num_inputs", C.c_int), ("output_ports", C.c_uint8 * 8), ("num_outputs", C.c_int),
                ("tests", OrTest * 32), ("num_tests", C.c_int)]


class OrState(C.Structure):
    """search_state_t (src/search_state.h:19-25), 546 bytes."""
    _fields_ = [("regs", (C.c_uint16 * 8) * 32), ("flags", C.c_uint8 * 32), ("live", C.c_uint8),
                ("depth", C.c_uint8)]


class OrOpts(C.Structure):
    _fields_ = [("use_branches", C.c_int), ("oep_mode", C.c_int), ("oep_bits", C.c_int),
                ("max_solutions", C.c_int), ("start_depth", C.c_int), ("max_depth", C.c_int),
                ("time_limit", C.c_double), ("print", C.c_int), ("shadow_exact", C.c_int)]


class OrResult(C.Structure):
    _fields_ = [("optimal_depth", C.c_int), ("num_solutions", C.c_int),
                ("solutions", C.POINTER(C.c_uint16 * 16)), ("solution_lens", C.POINTER(C.c_int)),
                ("tried", C.c_uint64 * 17), ("leaf_tried", C.c_uint64 * 17), ("pruned", C.c_uint64 * 17),
                ("ref_collisions", C.c_uint64 * 17), ("ref_saturations", C.c_uint64 * 17),
                ("cap_overflow", C.c_int), ("timed_out", C.c_int), ("seconds", C.c_double),
                ("depths_done", C.c_int)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.or_task_load.argtypes = [C.c_char_p, C.POINTER(OrTask)]
        L.or_task_parse.argtypes = [C.c_char_p, C.POINTER(OrTask)]
        L.or_task_evaluate.argtypes = [C.POINTER(OrTask), C.POINTER(C.c_uint16), C.c_int]
        L.or_disasm.argtypes = [C.c_uint16, C.c_char_p, C.c_size_t]
        L.or_state_root.argtypes = [C.POINTER(OrTask), C.POINTER(OrState)]
        L.or_state_step.argtypes = [C.POINTER(OrTask), C.POINTER(OrState), C.c_uint16, C.POINTER(OrState)]
        L.or_state_leaf.argtypes = [C.POINTER(OrTask), C.POINTER(OrState)]
        L.or_state_leaf.restype = C.c_uint8
        L.or_complete.argtypes = [C.POINTER(OrTask), C.POINTER(OrState), C.c_int, C.c_int, C.POINTER(C.c_uint16),
                                  C.POINTER(C.c_int), C.POINTER(C.c_uint64)]
        L.or_disasm.restype = None
        L.or_gen_candidates.argtypes = [C.POINTER(C.c_uint16), C.c_int, C.c_uint8, C.c_int, C.c_int, C.c_int]
        L.or_gen_candidates_branchless.argtypes = [C.POINTER(C.c_uint16), C.c_int, C.c_uint8]
        L.or_opts_default.argtypes = [C.POINTER(OrOpts)]
        L.or_opts_default.restype = None
        L.or_solve.argtypes = [C.POINTER(OrTask), C.POINTER(OrOpts), C.POINTER(OrResult)]
        L.or_result_free.argtypes = [C.POINTER(OrResult)]
        L.or_result_free.restype = None
        L.or_frontier.argtypes = [C.POINTER(OrTask), C.c_int, C.c_int, C.POINTER(C.c_uint16),
                                  C.POINTER(C.c_uint8), C.c_long]
        L.or_frontier.restype = C.c_long
        L.or_suffix_hits.argtypes = [C.POINTER(OrTask), C.c_int, C.POINTER(C.c_uint16), C.c_int, C.c_uint8,
                                     C.POINTER(C.c_uint16), C.c_long, C.POINTER(C.c_uint64)]
        L.or_suffix_hits.restype = C.c_long
        L.or_replay_state.argtypes = [C.POINTER(OrTask), C.POINTER(C.c_uint16), C.c_int, C.POINTER(C.c_uint16)]
        L.or_replay_state.restype = None
        L.or_alu.argtypes = [C.c_int, C.c_uint16, C.c_uint16, C.c_uint8, C.POINTER(C.c_uint8), C.POINTER(C.c_int)]
        L.or_alu.restype = C.c_uint16
        L.or_last_level.argtypes = [C.POINTER(OrTask), C.c_int, C.POINTER(C.c_uint16), C.POINTER(C.c_uint8),
                                    C.c_long, C.c_int, C.POINTER(C.c_uint32), C.c_long, C.POINTER(C.c_uint64)]
        L.or_last_level.restype = C.c_long
        L.or_live_after.argtypes = [C.c_uint8, C.c_uint16]
        L.or_live_after.restype = C.c_uint8
        _L = L
    return _L


def task_from_bytes(raw):
    """An oracle task from the raw bytes of a product Task (same layout)."""
    t = OrTask()
    C.memmove(C.byref(t), raw, C.sizeof(OrTask))
    return t


def load_task(path):
    t = OrTask()
    if lib().or_task_load(os.fsencode(path), C.byref(t)) < 0:
        raise IOError(path)
    return t


def parse_task(text):
    t = OrTask()
    lib().or_task_parse(text.encode(), C.byref(t))
    return t


def evaluate(task, prog):
    arr = (C.c_uint16 * len(prog))(*prog)
    return lib().or_task_evaluate(C.byref(task), arr, len(prog))


def disasm(insn):
    buf = C.create_string_buffer(64)
    lib().or_disasm(insn, buf, 64)
    return buf.value.decode()


def gen_candidates(live, pos, depth, branches=True):
    out = (C.c_uint16 * 4096)()
    n = lib().or_gen_candidates(out, 4096, live, pos, depth, int(branches))
    return list(out[:n])


def gen_candidates_branchless(live):
    out = (C.c_uint16 * 4096)()
    n = lib().or_gen_candidates_branchless(out, 4096, live)
    return list(out[:n])


def alu(op, a, b, flags=0):
    f = C.c_uint8()
    fault = C.c_int()
    v = lib().or_alu(op, a, b, flags, C.byref(f), C.byref(fault))
    return v, f.value, bool(fault.value)


def solve(task, max_depth, oep=OEP_EXACT, branches=True, max_solutions=100, start_depth=0,
          time_limit=0.0, shadow=False):
    o = OrOpts()
    lib().or_opts_default(C.byref(o))
    o.use_branches = int(branches)
    o.oep_mode = oep
    o.max_solutions = max_solutions
    o.start_depth = start_depth
    o.max_depth = max_depth
    o.time_limit = time_limit
    o.shadow_exact = int(shadow)
    r = OrResult()
    lib().or_solve(C.byref(task), C.byref(o), C.byref(r))
    sols = [list(r.solutions[i][:r.solution_lens[i]]) for i in range(r.num_solutions)]
    out = {
        "optimal_depth": r.optimal_depth, "solutions": sols, "num_solutions": r.num_solutions,
        "tried": list(r.tried), "leaf_tried": list(r.leaf_tried), "pruned": list(r.pruned),
        "ref_events": [a + b for a, b in zip(r.ref_collisions, r.ref_saturations)],
        "cap_overflow": bool(r.cap_overflow), "timed_out": bool(r.timed_out), "seconds": r.seconds,
        "depths_done": r.depths_done,
    }
    lib().or_result_free(C.byref(r))
    return out


def frontier(task, depth, split, cap=1 << 22):
    ins = (C.c_uint16 * (8 * cap))()
    live = (C.c_uint8 * cap)()
    n = lib().or_frontier(C.byref(task), depth, split, ins, live, cap)
    if n < 0:
        raise OverflowError("frontier larger than cap")
    return [(tuple(ins[i * 8:i * 8 + split]), live[i]) for i in range(n)]


def frontier_arrays(task, depth, split, cap=1 << 18, max_cap=1 << 28):
    """The frontier as flat ctypes arrays (8 kernel words per prefix, live
    masks) and its length: the form last_level takes, for large levels.
    The buffers start at cap prefixes and grow x8 (the DFS is re-run) up
    to max_cap."""
    while True:
        ins = (C.c_uint16 * (8 * cap))()
        live = (C.c_uint8 * cap)()
        n = lib().or_frontier(C.byref(task), depth, split, ins, live, cap)
        if n >= 0:
            return ins, live, n
        if cap >= max_cap:
            raise OverflowError("frontier larger than max_cap")
        cap *= 8


def last_level(task, depth, prefixes=None, arrays=None, threads=None, cap=1 << 16):
    """or_last_level: rows (prefix index, j, mask, cannot_pass, first) of the
    children at kernel position depth-1 whose state has a register equal to
    the target on every test, in DFS order, and the number of children.
    Give either prefixes [(insns, live)] or arrays (ins, live, n)."""
    if threads is None:
        threads = max(1, min(8, os.cpu_count() or 1))
    if arrays is None:
        n = len(prefixes)
        ins = (C.c_uint16 * (8 * max(n, 1)))()
        live = (C.c_uint8 * max(n, 1))()
        for i, (p, lv) in enumerate(prefixes):
            for q, v in enumerate(p):
                ins[i * 8 + q] = v
            live[i] = lv
    else:
        ins, live, n = arrays
    children = C.c_uint64()
    while True:
        rows = (C.c_uint32 * (4 * cap))()
        m = lib().or_last_level(C.byref(task), depth, ins, live, n, threads, rows, cap, C.byref(children))
        if m < 0:
            raise MemoryError("or_last_level")
        if m <= cap:
            break
        cap = m
    out = [(rows[4 * i], rows[4 * i + 1], rows[4 * i + 2], rows[4 * i + 3] & 1, (rows[4 * i + 3] >> 1) & 1)
           for i in range(m)]
    return out, children.value


def solve_at_depth(task, depth, threads=None):
    """The exact-OEP solutions of kernel depth `depth` alone (tools/enumerate.c
    IDDFS at target depth D, every solution kept), computed as the frontier
    at D-1 (the DFS with exact keys) followed by the last position on
    or_last_level: try_output on every first-visitor child, in DFS order.
    Returns (programs, frontier size, children at the last position)."""
    if depth == 0:
        r = solve(task, 0, OEP_EXACT, max_solutions=0, start_depth=0)
        return r["solutions"], 1, 0
    ins, live, n = frontier_arrays(task, depth, depth - 1)
    rows, children = last_level(task, depth, arrays=(ins, live, n), threads=threads)
    L = lib()
    nin = task.num_inputs
    sols = []
    cache = {}
    for pidx, j, mask, dead, first in rows:
        if not first or dead:
            continue
        if pidx not in cache:
            cache = {pidx: gen_candidates(live[pidx], depth - 1, depth)}
        c = cache[pidx][j]
        kern = list(ins[pidx * 8:pidx * 8 + depth - 1]) + [c]
        lv = L.or_live_after(live[pidx], c)
        prog = [(23 << 11) | ((q + 1) << 8) | task.input_ports[q] for q in range(nin)] + kern
        for r in range(1, 8):
            if not (lv >> r & 1 and mask >> r & 1):
                continue
            full = prog + [(24 << 11) | (r << 8) | task.output_ports[0], 1 << 11]
            if evaluate(task, full) == task.num_tests:
                sols.append(full)
    return sols, n, children


def suffix_hits(task, depth, prefix, split, live, cap=1 << 20):
    """Brute-force suffix hits below one prefix: rows (j_s..j_{D-1}, out_reg)."""
    k = depth - split
    rows = (C.c_uint16 * ((k + 1) * cap))()
    pre = (C.c_uint16 * 8)(*prefix)
    cands = C.c_uint64()
    n = lib().or_suffix_hits(C.byref(task), depth, pre, split, live, rows, cap, C.byref(cands))
    m = min(n, cap)
    return [tuple(rows[i * (k + 1):(i + 1) * (k + 1)]) for i in range(m)], n, cands.value


def replay_state(task, kernel):
    arr = (C.c_uint16 * max(len(kernel), 1))(*kernel)
    out = (C.c_uint16 * (9 * 32))()
    lib().or_replay_state(C.byref(task), arr, len(kernel), out)
    return [tuple(out[t * 9:(t + 1) * 9]) for t in range(task.num_tests)]


def state_root(task):
    st = OrState()
    lib().or_state_root(C.byref(task), C.byref(st))
    return st


def state_step(task, st, insn):
    """The child state, or None if the step is rejected."""
    out = OrState()
    return out if lib().or_state_step(C.byref(task), C.byref(st), insn, C.byref(out)) else None


def state_leaf(task, st):
    return lib().or_state_leaf(C.byref(task), C.byref(st))


def state_after(task, insns):
    """Root state stepped through insns (None if a step is rejected)."""
    st = state_root(task)
    for c in insns:
        st = state_step(task, st, c)
        if st is None:
            return None
    return st


def complete(task, st, max_depth, oep=OEP_EXACT):
    """search_exhaustive_complete: (found, insns, explored)."""
    ins = (C.c_uint16 * 16)()
    n = C.c_int()
    ex = C.c_uint64()
    rc = lib().or_complete(C.byref(task), C.byref(st), max_depth, oep, ins, C.byref(n), C.byref(ex))
    if rc < 0:
        raise MemoryError("or_complete")
    return bool(rc), list(ins[:n.value]), ex.value
