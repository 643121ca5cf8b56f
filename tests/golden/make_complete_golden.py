"""Writes the leaf-completion fixtures under tests/golden/complete/ (run in
the build container, where the reference checkout and oracle/_ref exist).

  tasks/<name>.json  task files of the reference's own dataset
                     (data/manifest.json, data/synthetic*/), copied as data
  cases.json         per case: task, the instructions stepped from the root
                     to the start state, max_depth, and what the reference's
                     search_exhaustive_complete (src/search_state.c:418-445,
                     compiled from the reference sources into
                     oracle/_ref/libpca.so by `make -C oracle ref`) returns:
                     found, the instruction list, states explored, seconds.

The cases follow the reference's own test (python/test_exhaustive.py:53-110):
root completion at the manifest's optimal depth d (found, length d), at d-1
(not found), the state after the first returned instruction (found within
d-1), plus seeded random start states one or two instructions below the
root with 1-3 instructions of budget, and a start state that already solves.

Usage: python tests/golden/make_complete_golden.py
"""
import ctypes as C
import json
import os
import random
import shutil
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
OUT = os.path.join(HERE, "complete")
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402

# tasks per (manifest optimal depth, inputs): the first ones in manifest order
PER_DEPTH = {(1, 1): 2, (1, 2): 2, (2, 1): 3, (2, 2): 3, (3, 1): 3, (3, 2): 3, (4, 1): 1, (4, 2): 1}
# random start states per task: (instructions below the root, budget)
RANDOM = [(1, 1), (1, 2), (2, 1), (2, 2), (1, 3)]
# two-input tasks: start states two instructions below the root with a
# budget of 3 (the reference loop's completion radius, python/bootstrap.py:331)
HEAVY = [(2, 3), (2, 3)]


def ref_lib():
    L = C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libpca.so"))
    L.task_load.argtypes = [C.c_char_p, C.c_void_p]
    L.search_ctx_init.argtypes = [C.c_void_p, C.c_void_p]
    L.search_state_init_root.argtypes = [C.c_void_p, C.c_void_p]
    L.search_state_step.argtypes = [C.c_void_p, C.c_void_p, C.c_uint16, C.c_void_p]
    L.search_exhaustive_complete.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_uint16),
                                             C.POINTER(C.c_int)]
    L.search_exhaustive_states_explored.restype = C.c_uint64
    L.search_gen_candidates_branchless.argtypes = [C.POINTER(C.c_uint16), C.c_int, C.c_void_p]
    return L


class Ref:
    def __init__(self, L, path):
        self.L = L
        self.spec = C.create_string_buffer(4096)
        assert L.task_load(path.encode(), self.spec) == 0, path
        self.ctx = C.create_string_buffer(8192)
        assert L.search_ctx_init(self.ctx, self.spec) == 0, path

    def state(self, insns):
        st = C.create_string_buffer(546)
        self.L.search_state_init_root(self.ctx, st)
        for c in insns:
            nx = C.create_string_buffer(546)
            if not self.L.search_state_step(self.ctx, st, c, nx):
                return None
            st = nx
        return st

    def cands(self, st):
        out = (C.c_uint16 * 4096)()
        n = self.L.search_gen_candidates_branchless(out, 4096, st)
        return list(out[:n])

    def complete(self, st, max_depth):
        ins = (C.c_uint16 * 16)()
        n = C.c_int()
        t0 = time.time()
        found = self.L.search_exhaustive_complete(self.ctx, st, max_depth, ins, C.byref(n))
        dt = time.time() - t0
        return bool(found), list(ins[:n.value]), int(self.L.search_exhaustive_states_explored()), dt


def main():
    L = ref_lib()
    with open(os.path.join(REF, "data", "manifest.json")) as f:
        manifest = json.load(f)
    picked = []
    for (d, ni), n in sorted(PER_DEPTH.items()):
        picked += [t for t in manifest["tasks"]
                   if t.get("solved") and t.get("optimal_depth") == d and t["num_inputs"] == ni][:n]
    os.makedirs(os.path.join(OUT, "tasks"), exist_ok=True)
    cases = []
    for t in picked:
        src = os.path.join(REF, t["path"])
        name = t["name"]
        shutil.copyfile(src, os.path.join(OUT, "tasks", name + ".json"))
        d = t["optimal_depth"]
        ref = Ref(L, src)

        def add(kind, path, max_depth):
            st = ref.state(path)
            found, ins, explored, dt = ref.complete(st, max_depth)
            cases.append({"task": name, "kind": kind, "optimal_depth": d, "path": path, "max_depth": max_depth,
                          "found": found, "insns": ins, "explored": explored, "seconds": round(dt, 4)})
            print(f"{name:28s} {kind:7s} path {len(path)} max {max_depth}: found {found} len {len(ins)} "
                  f"explored {explored} ({dt:.2f}s)", flush=True)
            return found, ins

        found, ins = add("root", [], d)
        # the manifest's depth is the optimum for d <= 3 (the reference's
        # test relies on it); deeper entries may be loop-recorded lengths
        assert found and (len(ins) == d or d > 3), (name, ins)
        d_found = len(ins)
        if d_found >= 2:
            add("below", [], d_found - 1)
            add("mid", ins[:1], d)
        add("solved", ins, d)
        rng = random.Random(sum(map(ord, name)))
        if d <= 3:
            for depth, budget in RANDOM:
                path = []
                for _ in range(depth):
                    path.append(rng.choice(ref.cands(ref.state(path))))
                add("random", path, depth + budget)
        if d == 3 and t["num_inputs"] == 2:
            for depth, budget in HEAVY:
                path = []
                for _ in range(depth):
                    path.append(rng.choice(ref.cands(ref.state(path))))
                add("heavy", path, depth + budget)
    with open(os.path.join(OUT, "cases.json"), "w") as f:
        json.dump({"source": "reference search_exhaustive_complete (oracle/_ref/libpca.so)", "cases": cases},
                  f, indent=0)
        f.write("\n")
    print(f"{len(cases)} cases")


if __name__ == "__main__":
    main()
