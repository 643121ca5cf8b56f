"""Writes the golden fixtures under tests/golden/ (run in the build
container, where the reference checkout and oracle/_ref exist).

  ref/<task>_<args>.txt   stdout of the reference binary oracle/_ref/enumerate
                          (built from the reference sources by
                          `make -C oracle ref`), wall-clock figures masked.
                          The oracle's --ref32 mode must reproduce these
                          byte for byte.
  exact/<task>_d<D>.json  the oracle's exact-key IDDFS (every solution kept):
                          optimal depth, the canonical program list and the
                          per-depth tried / pruned counters.  For runs with no
                          32-bit table events these equal the reference's
                          output (checked when they are written).

Usage: python tests/golden/make_golden.py [ref|exact|all] [names...]
"""
import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

REF_RUNS = [
    ("negate", ["-d", "4"]), ("double", ["-d", "4"]), ("square", ["-d", "4"]), ("add", ["-d", "4"]),
    ("abs", ["-d", "4"]), ("relu", ["-d", "4"]), ("sign", ["-d", "2"]), ("clamp", ["-d", "2"]),
    ("popcount", ["-d", "2"]), ("fib", ["-d", "2"]), ("gcd", ["-d", "2"]), ("collatz", ["-d", "2"]),
    ("max", ["-d", "2"]), ("is_power2", ["-d", "2"]), ("saturating_add", ["-d", "2", "--no-branch"]),
    ("negate", ["-d", "3", "--no-branch"]), ("add", ["-d", "3", "--no-branch"]),
    ("abs", ["-d", "3", "--no-branch"]), ("double", ["-d", "3", "--no-oep"]),
    ("square", ["-d", "2", "--no-branch"]),
]

EXACT_RUNS = [("negate", 3), ("double", 3), ("square", 3), ("add", 3), ("abs", 3), ("relu", 3),
              ("sign", 3), ("clamp", 2), ("popcount", 2), ("is_power2", 2), ("max", 2),
              ("negate_16", 5), ("is_power2_32", 2)]


def mask(text):
    text = re.sub(r"\([0-9.]+ M/sec\)", "(X M/sec)", text)
    text = re.sub(r"elapsed: [0-9.]+ sec", "elapsed: X sec", text)
    return re.sub(r"in [0-9.]+ sec", "in X sec", text)


def task_path(name):
    return os.path.join(HERE, "tasks", name + ".json")


def write_generated_tasks():
    from pca_amd import taskgen
    for name, fn in (("negate_16", taskgen.negate_16), ("is_power2_32", taskgen.is_power2_32)):
        with open(task_path(name), "w") as f:
            f.write(fn(42).to_json() + "\n")


def make_ref(names):
    ref = os.path.join(ROOT, "oracle", "_ref", "enumerate")
    os.makedirs(os.path.join(HERE, "ref"), exist_ok=True)
    for name, args in REF_RUNS:
        if names and name not in names:
            continue
        out = subprocess.run([ref, task_path(name)] + args, capture_output=True, text=True).stdout
        tag = name + "_" + "_".join(a.strip("-") for a in args)
        with open(os.path.join(HERE, "ref", tag + ".txt"), "w") as f:
            f.write(mask(out))
        print("ref", tag, flush=True)


def make_exact(names):
    os.makedirs(os.path.join(HERE, "exact"), exist_ok=True)
    for name, d in EXACT_RUNS:
        if names and name not in names:
            continue
        t = O.load_task(task_path(name))
        r = O.solve(t, d, O.OEP_EXACT, max_solutions=0)
        shadow = O.solve(t, d, O.OEP_REF32, max_solutions=0, shadow=True)
        rec = {"task": name, "max_depth": d, "optimal_depth": r["optimal_depth"],
               "num_solutions": r["num_solutions"], "solutions": r["solutions"],
               "tried": r["tried"][:d + 1], "pruned": r["pruned"][:d + 1],
               "ref32_events": shadow["ref_events"][:d + 1],
               "equals_ref32": shadow["solutions"] == r["solutions"] and
               shadow["optimal_depth"] == r["optimal_depth"]}
        with open(os.path.join(HERE, "exact", f"{name}_d{d}.json"), "w") as f:
            json.dump(rec, f)
        print("exact", name, d, r["optimal_depth"], r["num_solutions"], "events", sum(rec["ref32_events"]),
              flush=True)


if __name__ == "__main__" and (len(sys.argv) < 2 or sys.argv[1] != "alu"):
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    names = sys.argv[2:]
    write_generated_tasks()
    if what in ("ref", "all"):
        make_ref(names)
    if what in ("exact", "all"):
        make_exact(names)


ALU_OPS = {"ADD": 2, "SUB": 3, "MUL": 4, "DIV": 5, "MOD": 6, "AND": 7, "OR": 8, "XOR": 9, "SHL": 10,
           "SHR": 11, "ASR": 12, "CMP": 13}


def make_alu_digests():
    """alu_digests.json: the oracle's exhaustive digest of every ALU opcode
    over all 2^32 operand pairs (see or_alu_digest)."""
    import ctypes as C
    L = O.lib()
    L.or_alu_digest.restype = C.c_uint64
    L.or_alu_digest.argtypes = [C.c_int, C.c_uint32, C.c_uint32]
    out = {}
    for name, op in ALU_OPS.items():
        out[name] = {"op": op, "full": str(L.or_alu_digest(op, 0, 65536)),
                     "b_0_256": str(L.or_alu_digest(op, 0, 256)),
                     "b_65280_65536": str(L.or_alu_digest(op, 65280, 65536))}
        print("alu", name, flush=True)
    with open(os.path.join(HERE, "alu_digests.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__" and len(sys.argv) > 1 and sys.argv[1] == "alu":
    make_alu_digests()
