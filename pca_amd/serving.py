"""Continuous-batching serving over the C ABI (pcagpu_server_*).

``Server`` wraps the library's request loop: requests (task + depth budget)
are admitted into resident device slots as they free up, every active task
contributes its next prefix batch to one shared persistent scan launch, and
each result is the task's exact canonical answer.

``poisson_run`` replays a Poisson arrival schedule against the server, and
``sequential_run`` replays the same schedule through the plain
one-task-at-a-time GPU path (``Solver.solve`` per request, in arrival order),
so the two latency distributions are comparable.
"""
import ctypes as C
import random
import time

from ._lib import (PcaGpuError, Task, Solution, lib, ServerOpts, Request, ServerStats, SERVER_KEEP)
from .solver import Solver, _sol_tuple

QUEUED, ACTIVE, DONE, FAILED = 0, 1, 2, 3


class Server:
    """One continuous-batching server (``pcagpu_server_*``).

    ``max_slots`` / ``hbm_budget`` bound the resident tasks; ``warm_slots``
    slots are opened now (0: none) so early requests allocate nothing; a
    task adds at least ``prefixes_per_item`` prefixes to a step, and
    ``step_prefixes`` split over the active tasks when that is more."""

    def __init__(self, device=0, max_slots=64, hbm_budget=200 << 30, prefixes_per_item=1 << 14, warm_slots=4,
                 step_prefixes=1 << 18, hit_capacity=0, fail_step=0):
        """hit_capacity: device hit records per slot (0: 65536) before a batch
        is rescanned alone with a larger buffer; fail_step (testing): the
        n-th step fails after staging, as a failed allocation would."""
        o = ServerOpts(device=device, max_slots=max_slots, hbm_budget=hbm_budget,
                       prefixes_per_item=prefixes_per_item,
                       warm_slots=0xFFFFFFFF if warm_slots == 0 else warm_slots, step_prefixes=step_prefixes,
                       hit_capacity=hit_capacity, fail_step=fail_step)
        h = C.c_void_p()
        rc = lib().pcagpu_server_open(C.byref(o), C.byref(h))
        if rc < 0:
            raise PcaGpuError(rc, "pcagpu_server_open")
        self._h = h

    def close(self):
        if self._h:
            lib().pcagpu_server_close(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def clock(self):
        return lib().pcagpu_server_clock(self._h)

    def submit(self, task, max_depth):
        rid = C.c_uint64()
        rc = lib().pcagpu_server_submit(self._h, C.byref(task), max_depth, C.byref(rid))
        if rc < 0:
            raise PcaGpuError(rc, "server_submit")
        return rid.value

    def step(self):
        rc = lib().pcagpu_server_step(self._h)
        if rc < 0:
            raise PcaGpuError(rc, "server_step")
        return rc

    def poll(self, cap=256):
        buf = (Request * cap)()
        n = lib().pcagpu_server_poll(self._h, buf, cap)
        if n < 0:
            raise PcaGpuError(n, "server_poll")
        return [{"id": r.id, "status": r.status, "optimal_depth": r.optimal_depth,
                 "num_solutions": r.num_solutions, "max_depth": r.max_depth, "t_submit": r.t_submit,
                 "t_admit": r.t_admit, "t_done": r.t_done} for r in buf[:n]]

    def solutions(self, rid, cap=4096):
        buf = (Solution * cap)()
        n = lib().pcagpu_server_solutions(self._h, rid, buf, cap)
        if n < 0:
            raise PcaGpuError(n, "server_solutions")
        return [_sol_tuple(buf[i]) for i in range(min(n, cap, SERVER_KEEP))], n

    def stats(self):
        st = ServerStats()
        lib().pcagpu_server_get_stats(self._h, C.byref(st))
        return st.as_dict()

    def drain(self, timeout=600.0):
        """Steps until every submitted request is done; returns polled records."""
        out = []
        t0 = time.perf_counter()
        while True:
            pending = self.step()
            out += self.poll()
            if pending == 0:
                out += self.poll()
                return out
            if time.perf_counter() - t0 > timeout:
                raise TimeoutError("server drain")


def poisson_schedule(n, rate, seed=0):
    rng = random.Random(seed)
    t, out = 0.0, []
    for _ in range(n):
        t += rng.expovariate(rate)
        out.append(t)
    return out


def _programs(sols):
    return [x[0] for x in sols]


def poisson_run(tasks, budgets, arrivals, warm=(), **server_kw):
    """Replays arrivals (seconds) against a Server; returns per-request
    (arrival, done, optimal_depth, num_solutions, programs) -- programs: the
    canonical program list, the first SERVER_KEEP of them -- and the server
    stats.  Programs are read back after the run, off the clock.  `warm`:
    (task, budget) requests served before the clock starts (a long-running
    server's buffers already grown)."""
    with Server(**server_kw) as sv:
        for t, b in warm:
            sv.submit(t, b)
        sv.drain()
        ids = {}
        recs = {}
        i = 0
        t0 = time.perf_counter()
        while len(recs) < len(tasks):
            now = time.perf_counter() - t0
            while i < len(tasks) and arrivals[i] <= now:
                ids[sv.submit(tasks[i], budgets[i])] = i
                i += 1
            pending = sv.step()
            now = time.perf_counter() - t0
            for r in sv.poll():
                recs[ids[r["id"]]] = [arrivals[ids[r["id"]]], now, r["optimal_depth"], r["num_solutions"], r["id"],
                                      r["status"]]
            if pending == 0 and i < len(tasks):
                time.sleep(max(0.0, min(arrivals[i] - (time.perf_counter() - t0), 0.001)))
        out = []
        for k in range(len(tasks)):
            arr, done, od, ns, rid, status = recs[k]
            progs = _programs(sv.solutions(rid)[0]) if status == DONE else None
            out.append((arr, done, od, ns, progs))
        return out, sv.stats()


def sequential_run(tasks, budgets, arrivals, warm=True, warm_tasks=()):
    """The same arrivals through the plain one-task-at-a-time GPU path
    (pcagpu_solve).  warm: one resident handle re-tasked per request, opened
    before the first arrival (allocation off the clock, as the server's
    warm slots), after solving warm_tasks [(task, budget)]; otherwise a
    handle opened per request."""
    recs = []
    sv = Solver(tasks[0]) if warm and tasks else None
    try:
        if sv is not None:
            for t, b in warm_tasks:
                sv.retask(t)
                sv.solve(b, cap=1)
        t0 = time.perf_counter()
        for task, budget, arr in zip(tasks, budgets, arrivals):
            now = time.perf_counter() - t0
            if now < arr:
                time.sleep(arr - now)
            if sv is not None:
                sv.retask(task)
                r = sv.solve(budget, cap=SERVER_KEEP)
            else:
                with Solver(task) as s:
                    r = s.solve(budget, cap=SERVER_KEEP)
            recs.append((arr, time.perf_counter() - t0, r.optimal_depth, r.num_solutions, _programs(r.solutions)))
    finally:
        if sv is not None:
            sv.close()
    return recs


def suite_run(tasks, budget, timeout=1200.0, **server_kw):
    """A batched suite: every task submitted at once, each server step one
    persistent launch over all active tasks.  Returns per-task
    (optimal_depth, num_solutions, programs), the wall time from the first
    submit to the last result, the server stats (with ``open_s``, the time
    ``pcagpu_server_open`` took, warm slots included)."""
    t_open = time.perf_counter()
    with Server(**server_kw) as sv:
        t0 = time.perf_counter()
        ids = {sv.submit(t, budget): i for i, t in enumerate(tasks)}
        done = [None] * len(tasks)
        for r in sv.drain(timeout):
            done[ids[r["id"]]] = r
        wall = time.perf_counter() - t0
        st = sv.stats()
        st["open_s"] = t0 - t_open
        out = [(r["optimal_depth"], r["num_solutions"],
                _programs(sv.solutions(r["id"])[0]) if r["status"] == DONE else None) for r in done]
        return out, wall, st


def sequential_suite(tasks, budget, warm=True):
    """The same suite one task at a time through Solver.solve (one handle
    re-tasked per task when warm, opened before the clock starts)."""
    out = []
    sv = Solver(tasks[0]) if warm and tasks else None
    try:
        t0 = time.perf_counter()
        for t in tasks:
            if sv is not None:
                sv.retask(t)
                r = sv.solve(budget, cap=SERVER_KEEP)
            else:
                with Solver(t) as s:
                    r = s.solve(budget, cap=SERVER_KEEP)
            out.append((r.optimal_depth, r.num_solutions, _programs(r.solutions)))
        return out, time.perf_counter() - t0
    finally:
        if sv is not None:
            sv.close()


def latency_stats(recs):
    lat = sorted(r[1] - r[0] for r in recs)
    if not lat:
        return {}

    def pct(p):
        return lat[min(len(lat) - 1, int(round(p / 100.0 * (len(lat) - 1))))]
    return {"p50_s": pct(50), "p99_s": pct(99), "mean_s": sum(lat) / len(lat), "n": len(lat)}
