"""ctypes binding of include/pcagpu.h (the library's C ABI).

Struct layouts mirror the header byte for byte; ``Task`` is byte-compatible
with the reference's ``task_spec_t`` (tasks/spec.h:20-34) and ``Prefix`` with
``prefix_t`` (gpu/hybrid.cu:56-60).
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpcagpu.so")

MAX_TESTS = 32
MAX_PORTS = 8
MAX_KDEPTH = 8
MAX_PROG = 16
SERVER_KEEP = 4096   # PCAGPU_SERVER_KEEP: solutions a server keeps per request

ERRORS = {
    -1: "EINVAL (bad argument or unsupported task)",
    -2: "ENOMEM (allocation failed)",
    -3: "EHIP (HIP runtime call failed)",
    -4: "ESTATE (call out of order)",
    -5: "EOVERFLOW (fixed-capacity buffer overflowed)",
    -6: "ECOLLISION (64-bit OEP key shared by two distinct states)",
    -7: "ENODEV (no GPU visible)",
    -8: "EIO (task file unreadable)",
    -9: "EVERIFY (a device result failed host re-verification)",
}


class PcaGpuError(RuntimeError):
    def __init__(self, code, where=""):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)}")


class TestCase(C.Structure):
    _fields_ = [("inp", C.c_uint16 * MAX_PORTS), ("out", C.c_uint16 * MAX_PORTS)]


class Task(C.Structure):
    _fields_ = [
        ("name", C.c_char * 64),
        ("max_cycles", C.c_uint64),
        ("input_ports", C.c_uint8 * MAX_PORTS),
        ("num_inputs", C.c_int),
        ("output_ports", C.c_uint8 * MAX_PORTS),
        ("num_outputs", C.c_int),
        ("tests", TestCase * MAX_TESTS),
        ("num_tests", C.c_int),
    ]

    @classmethod
    def load(cls, path):
        t = cls()
        rc = lib().pcagpu_task_load(os.fsencode(path), C.byref(t))
        if rc < 0:
            raise PcaGpuError(rc, f"task_load({path})")
        return t

    @classmethod
    def parse(cls, text):
        t = cls()
        rc = lib().pcagpu_task_parse(text.encode(), C.byref(t))
        if rc < 0:
            raise PcaGpuError(rc, "task_parse")
        return t

    @classmethod
    def from_io(cls, name, inputs, outputs, max_cycles=100, input_ports=None, output_ports=(0,)):
        """Builds a task from per-test input tuples and single outputs."""
        t = cls()
        t.name = name.encode()[:63]
        t.max_cycles = max_cycles
        n_in = len(inputs[0]) if inputs else 0
        ports = list(input_ports) if input_ports is not None else list(range(n_in))
        t.num_inputs = n_in
        for i, p in enumerate(ports):
            t.input_ports[i] = p
        t.num_outputs = len(output_ports)
        for i, p in enumerate(output_ports):
            t.output_ports[i] = p
        if len(inputs) > MAX_TESTS:
            raise ValueError("at most 32 tests")
        t.num_tests = len(inputs)
        for k, (ins, out) in enumerate(zip(inputs, outputs)):
            for i, v in enumerate(ins):
                t.tests[k].inp[i] = v & 0xFFFF
            t.tests[k].out[0] = out & 0xFFFF
        return t

    def to_json(self):
        import json
        return json.dumps({
            "name": self.name.decode(), "max_cycles": int(self.max_cycles),
            "input_ports": [self.input_ports[i] for i in range(self.num_inputs)],
            "output_ports": [self.output_ports[i] for i in range(self.num_outputs)],
            "tests": [{"in": [self.tests[k].inp[i] for i in range(self.num_inputs)],
                       "out": [self.tests[k].out[i] for i in range(self.num_outputs)]}
                      for k in range(self.num_tests)]}, indent=1)

    def evaluate(self, prog):
        arr = (C.c_uint16 * len(prog))(*prog)
        return lib().pcagpu_task_evaluate(C.byref(self), arr, len(prog))


class Prefix(C.Structure):
    _pack_ = 1
    _fields_ = [("insns", C.c_uint16 * MAX_KDEPTH), ("num_insns", C.c_uint8), ("live", C.c_uint8)]


class Solution(C.Structure):
    _fields_ = [
        ("prefix_rank", C.c_uint64),
        ("prog", C.c_uint16 * MAX_PROG),
        ("len", C.c_uint8),
        ("out_reg", C.c_uint8),
        ("kernel_len", C.c_uint8),
        ("cpu_verified", C.c_uint8),
        ("cand", C.c_uint16 * MAX_KDEPTH),
    ]

    def program(self):
        return [self.prog[i] for i in range(self.len)]


class SearchState(C.Structure):
    """search_state_t (src/search_state.h:19-25), 546 bytes: per-test
    registers and flags, live mask, kernel instructions so far."""
    _fields_ = [("regs", (C.c_uint16 * 8) * MAX_TESTS), ("flags", C.c_uint8 * MAX_TESTS), ("live", C.c_uint8),
                ("depth", C.c_uint8)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "candidates", "suffixes", "tests_evaluated", "prefixes", "raw_hits", "kernel_ns",
        "launches", "frontier_nodes", "frontier_kept", "expand_ns", "oep_collisions", "parents", "scan_tile",
        "tests_executed")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


class Opts(C.Structure):
    _fields_ = [("device", C.c_int), ("batch_capacity", C.c_uint32), ("hit_capacity", C.c_uint32),
                ("waves_per_cu", C.c_uint32), ("fixed_test_order", C.c_uint32), ("scan_tile", C.c_uint32),
                ("eager_flush", C.c_uint32), ("test_flags", C.c_uint32)]


class SolveOpts(C.Structure):
    _fields_ = [("max_depth", C.c_int), ("start_depth", C.c_int), ("split", C.c_int),
                ("producer", C.c_int), ("verbose", C.c_int), ("time_limit", C.c_double)]


class SolveRes(C.Structure):
    _fields_ = [("optimal_depth", C.c_int), ("num_solutions", C.c_long), ("depths_done", C.c_int),
                ("timed_out", C.c_int), ("seconds", C.c_double),
                ("candidates", C.c_uint64 * (MAX_KDEPTH + 1)),
                ("prefixes", C.c_uint64 * (MAX_KDEPTH + 1)),
                ("prefix_nodes", C.c_uint64 * (MAX_KDEPTH + 1)),
                ("prefix_pruned", C.c_uint64 * (MAX_KDEPTH + 1)),
                ("depth_seconds", C.c_double * (MAX_KDEPTH + 1))]


class ServerOpts(C.Structure):
    _fields_ = [("device", C.c_int), ("max_slots", C.c_uint32), ("hbm_budget", C.c_uint64),
                ("prefixes_per_item", C.c_uint32), ("warm_slots", C.c_uint32), ("step_prefixes", C.c_uint32),
                ("hit_capacity", C.c_uint32), ("fail_step", C.c_uint32)]


class Request(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("optimal_depth", C.c_int),
                ("num_solutions", C.c_long), ("max_depth", C.c_int), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double), ("first", Solution)]


class ServerStats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in (
        "submitted", "admitted", "completed", "failed", "launches", "items", "kernel_ns", "candidates",
        "suffixes", "slots", "slots_opened", "device_bytes", "admit_ns", "stage_ns", "wait_ns", "resolve_ns",
        "rescans")]

    def as_dict(self):
        return {n: int(getattr(self, n)) for n, _ in self._fields_}


_LIB = None

# (name, restype, argtypes) for every symbol include/*.h declares
SIGNATURES = [
    ("pcagpu_task_load", C.c_int, [C.c_char_p, C.POINTER(Task)]),
    ("pcagpu_task_parse", C.c_int, [C.c_char_p, C.POINTER(Task)]),
    ("pcagpu_task_evaluate", C.c_int, [C.POINTER(Task), C.POINTER(C.c_uint16), C.c_int]),
    ("pcagpu_disasm", None, [C.c_uint16, C.c_char_p, C.c_size_t]),
    ("pcagpu_gen_candidates", C.c_int, [C.POINTER(C.c_uint16), C.c_int, C.c_uint8, C.c_int, C.c_int]),
    ("pcagpu_state_records", C.c_int, [C.POINTER(Task), C.POINTER(C.c_uint16), C.POINTER(C.c_uint8), C.c_int,
                                       C.POINTER(C.c_uint16), C.POINTER(C.c_uint8)]),
    ("pcagpu_open", C.c_int, [C.POINTER(Task), C.POINTER(Opts), C.POINTER(C.c_void_p)]),
    ("pcagpu_close", None, [C.c_void_p]),
    ("pcagpu_retask", C.c_int, [C.c_void_p, C.POINTER(Task)]),
    ("pcagpu_raw_hits_noalloc", C.c_long, [C.POINTER(Task), C.c_int, C.c_int, C.POINTER(Prefix), C.c_uint32,
                                           C.POINTER(C.c_uint16), C.c_long]),
    ("pcagpu_device_name", C.c_char_p, [C.c_void_p]),
    ("pcagpu_begin_depth", C.c_int, [C.c_void_p, C.c_int, C.c_int]),
    ("pcagpu_submit", C.c_int, [C.c_void_p, C.POINTER(Prefix), C.c_uint32, C.c_uint64]),
    ("pcagpu_flush", C.c_int, [C.c_void_p]),
    ("pcagpu_end_depth", C.c_long, [C.c_void_p, C.POINTER(Solution), C.c_long]),
    ("pcagpu_raw_hits", C.c_long, [C.c_void_p, C.POINTER(C.c_uint16), C.c_long]),
    ("pcagpu_alu_digest", C.c_int, [C.c_void_p, C.c_int, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("pcagpu_get_stats", C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    ("pcagpu_reset_stats", None, [C.c_void_p]),
    ("pcagpu_frontier", C.c_long, [C.c_void_p, C.c_int, C.c_int, C.POINTER(Prefix), C.c_long]),
    ("pcagpu_host_frontier", C.c_long, [C.POINTER(Task), C.c_int, C.c_int, C.POINTER(Prefix), C.c_long,
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("pcagpu_solve", C.c_int, [C.c_void_p, C.POINTER(SolveOpts), C.POINTER(SolveRes),
                               C.POINTER(Solution), C.c_long]),
    ("pcagpu_state_init_root", None, [C.POINTER(Task), C.POINTER(SearchState)]),
    ("pcagpu_state_step", C.c_int, [C.POINTER(Task), C.POINTER(SearchState), C.c_uint16, C.POINTER(SearchState)]),
    ("pcagpu_state_leaf_outputs", C.c_uint8, [C.POINTER(Task), C.POINTER(SearchState)]),
    ("pcagpu_gen_candidates_branchless", C.c_int, [C.POINTER(C.c_uint16), C.c_int, C.c_uint8]),
    ("pcagpu_complete", C.c_int, [C.c_void_p, C.POINTER(SearchState), C.c_int, C.POINTER(C.c_uint16),
                                  C.POINTER(C.c_int)]),
    ("pcagpu_complete_states_explored", C.c_uint64, [C.c_void_p]),
    ("pcagpu_server_open", C.c_int, [C.POINTER(ServerOpts), C.POINTER(C.c_void_p)]),
    ("pcagpu_server_close", None, [C.c_void_p]),
    ("pcagpu_server_submit", C.c_int, [C.c_void_p, C.POINTER(Task), C.c_int, C.POINTER(C.c_uint64)]),
    ("pcagpu_server_step", C.c_int, [C.c_void_p]),
    ("pcagpu_server_poll", C.c_int, [C.c_void_p, C.POINTER(Request), C.c_int]),
    ("pcagpu_server_solutions", C.c_long, [C.c_void_p, C.c_uint64, C.POINTER(Solution), C.c_long]),
    ("pcagpu_server_get_stats", C.c_int, [C.c_void_p, C.POINTER(ServerStats)]),
    ("pcagpu_server_clock", C.c_double, [C.c_void_p]),
    # include/pcavalue.h: device pointers and streams as integers
    ("pcav_gemm", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_int,
                            C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p]),
    ("pcav_features", C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pcav_relu_backward", C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pcav_pool_forward", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pcav_pool_backward", C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pcav_colsum", C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    ("pcav_bce", C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pcav_reach_dist_loss", C.c_int, [C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_float,
                                        C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pcav_adam", C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                            C.c_float, C.c_float, C.c_float, C.c_int, C.c_void_p]),
    ("pcav_to_bf16", C.c_int, [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("pcav_malloc", C.c_void_p, [C.c_int, C.c_size_t]),
    ("pcav_free", None, [C.c_void_p]),
    ("pcav_copy_to_device", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("pcav_copy_to_host", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t]),
    ("pcav_clear", C.c_int, [C.c_void_p, C.c_size_t]),
    ("pcav_synchronize", C.c_int, []),
]


def load_library(path=LIB_PATH):
    """Loads the native library.  Raises if it has not been built: there is
    no Python fallback for any part of the search path."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise PcaGpuError(-3, f"native library missing at {path}; run `python -m pca_amd.build`")
    L = C.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def lib():
    return load_library()


def disasm(insn):
    buf = C.create_string_buffer(64)
    lib().pcagpu_disasm(insn, buf, 64)
    return buf.value.decode()


def gen_candidates(live, pos, kernel_depth):
    out = (C.c_uint16 * 4096)()
    n = lib().pcagpu_gen_candidates(out, 4096, live, pos, kernel_depth)
    return list(out[:n])


def alu_digest(op, b_lo, b_hi, solver=None):
    """Exhaustive ALU digest, on the host (solver None) or the solver's GPU."""
    out = C.c_uint64()
    rc = lib().pcagpu_alu_digest(solver._h if solver is not None else None, op, b_lo, b_hi, C.byref(out))
    if rc < 0:
        raise PcaGpuError(rc, "alu_digest")
    return out.value


def host_frontier(task, target_depth, split, cap=1 << 22):
    """CPU prefix producer (exact-OEP DFS to the split); returns
    (prefixes as (insns tuple, live) list, nodes, pruned)."""
    buf = (Prefix * cap)()
    nodes = C.c_uint64()
    pruned = C.c_uint64()
    n = lib().pcagpu_host_frontier(C.byref(task), target_depth, split, buf, cap, C.byref(nodes), C.byref(pruned))
    if n < 0:
        raise PcaGpuError(n, "host_frontier")
    out = [(tuple(buf[i].insns[:split]), buf[i].live) for i in range(n)]
    return out, nodes.value, pruned.value
