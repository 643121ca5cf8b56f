"""MI355X-native GPU search path of the PCA-16 superoptimizer.

The product is the native library ``pca_amd/libpcagpu.so`` (HIP kernels for
gfx950 behind the C ABI in ``include/pcagpu.h``).  This package is the thin
Python host over that ABI, mirroring the reference's solver interface:

    from pca_amd import Task, Solver
    task = Task.load("tasks/negate.json")
    with Solver(task) as s:
        result = s.solve(max_depth=4)

Nothing here computes a search result in Python: every candidate is
evaluated by the device kernels, and the package raises if the native
library or a GPU is missing.
"""
from ._lib import (PcaGpuError, Task, Prefix, Solution, Stats, lib, load_library,
                   disasm, gen_candidates, host_frontier)
from .solver import Solver, SolveResult, raw_hits_noalloc
from . import search_state

__all__ = ["PcaGpuError", "Task", "Prefix", "Solution", "Stats", "lib", "load_library",
           "disasm", "gen_candidates", "host_frontier", "Solver", "SolveResult", "raw_hits_noalloc",
           "search_state"]
