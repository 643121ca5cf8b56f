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

  exact/<task>_d<D>.json  (deep / gen) depths the IDDFS cannot finish in
                          minutes, solved one depth at a time: the exact-key
                          DFS to D-1, then the last position brute force on
                          the VM (or_last_level); gen_* tasks are generated
                          ones with 24-32 tests and an optimum of 2-3.

Usage: python tests/golden/make_golden.py [ref|exact|all|deep|gen|alu] [names...]
"""
import json
import os
import re
import subprocess
import sys
import time

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


# Deep fixtures: depths the IDDFS of the oracle cannot finish in minutes.
# Every depth d below the optimum is solved alone (frontier at d-1, last
# position on or_last_level) and must be empty; the optimum's programs come
# from the same path.  Where the full IDDFS finishes (max at 3), it is run
# as well and must agree.
DEEP_RUNS = [("max", 3, True), ("is_power2", 4, False), ("is_power2_32", 4, False),
             ("saturating_add", 4, False)]

# Generated tasks of 24-32 tests whose optimum is 2 or 3 instructions: the
# config-2 test count with solutions, drawn from the chain and suite
# generators (pca_amd/taskgen.py), seed 2024.
GEN_SEED = 2024
GEN_WANT = 6


def _deep_record(name, t, maxd, check_iddfs):
    per_depth = []
    t0 = time.time()
    sols = []
    for d in range(maxd + 1):
        s0 = time.time()
        sols, nf, ch = O.solve_at_depth(t, d)
        per_depth.append({"depth": d, "frontier": nf, "last_children": ch, "solutions": len(sols),
                          "seconds": round(time.time() - s0, 2)})
        print(f"  {name} d{d}: frontier {nf}, children {ch}, {len(sols)} solutions, "
              f"{time.time() - s0:.1f}s", flush=True)
        if sols:
            break
    rec = {"task": name, "max_depth": maxd, "optimal_depth": d if sols else -1,
           "num_solutions": len(sols), "solutions": sols, "per_depth": per_depth,
           "method": "per depth: exact-key DFS to D-1 (or_frontier), last position by or_last_level",
           "seconds": round(time.time() - t0, 1)}
    if check_iddfs:
        r = O.solve(t, maxd, O.OEP_EXACT, max_solutions=0)
        assert r["optimal_depth"] == rec["optimal_depth"] and r["solutions"] == sols, name
        rec["checked_against_iddfs"] = True
        rec["tried"] = r["tried"][:maxd + 1]
        rec["pruned"] = r["pruned"][:maxd + 1]
    return rec


def make_deep(names):
    os.makedirs(os.path.join(HERE, "exact"), exist_ok=True)
    for name, d, check in DEEP_RUNS:
        if names and name not in names:
            continue
        rec = _deep_record(name, O.load_task(task_path(name)), d, check)
        with open(os.path.join(HERE, "exact", f"{name}_d{d}.json"), "w") as f:
            json.dump(rec, f)
        print("deep", name, d, rec["optimal_depth"], rec["num_solutions"], flush=True)


def make_generated(names):
    """gen_*: tasks with 24-32 tests and an optimum of 2-3 instructions."""
    from pca_amd import taskgen
    found = 0
    for i in range(400):
        if found >= GEN_WANT:
            break
        task = taskgen.chain_task(GEN_SEED, i, 3) if i % 2 == 0 else taskgen.random_task(GEN_SEED, i)
        if task.num_tests < 24:
            continue
        name = f"gen_{'chain' if i % 2 == 0 else 'suite'}_{i:03d}"
        ot = O.task_from_bytes(bytes(task))
        dstar = -1
        for d in range(4):
            sols, _, _ = O.solve_at_depth(ot, d)
            if sols:
                dstar = d
                break
        if dstar not in (2, 3):
            continue
        found += 1
        if names and name not in names:
            continue
        with open(task_path(name), "w") as f:
            f.write(task.to_json() + "\n")
        rec = _deep_record(name, O.load_task(task_path(name)), dstar, dstar == 2)
        with open(os.path.join(HERE, "exact", f"{name}_d{dstar}.json"), "w") as f:
            json.dump(rec, f)
        print("gen", name, "tests", task.num_tests, "inputs", task.num_inputs, "D*", dstar,
              rec["num_solutions"], flush=True)


if __name__ == "__main__" and len(sys.argv) > 1 and sys.argv[1] in ("deep", "gen"):
    (make_deep if sys.argv[1] == "deep" else make_generated)(sys.argv[2:])
    sys.exit(0)


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
