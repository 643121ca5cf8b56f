"""Value network that orders the guided search, trained and run on the GPU
with this package's BF16-MFMA kernels (include/pcavalue.h).

Network (per search state = a kernel prefix of a task):
  per test t < 32: 180 feature bits (2 x 16 input bits, 16 target bits,
  8 x 16 register bits, 4 flag bits: the reference's per-test layout,
  python/search_eval.py:206-251) -> 512 -> 256 (ReLU, shared weights);
  mean over the task's tests; concat a 16-wide embedding of the depth;
  -> 256 (ReLU) -> 10 outputs: a reach logit ("a solution is reachable from
  here within the depth budget") and logits over the remaining steps 0..8
  to the nearest solution (distance-to-solution), the two heads of the
  reference's model and loss (python/model.py:41-44, python/train.py:44-60:
  BCE on reach + 0.5 x cross-entropy on the distance of the reachable
  states).  296,594 parameters.  Weights are FP32 masters with BF16
  copies for the MFMA GEMMs, FP32 accumulation; activations are stored as
  BF16.  Device memory comes from the library's own allocator (one HIP
  runtime in the process) and every arithmetic step is a pcav_* kernel.

The network only orders candidates.  guided_search() returns a program
that the CPU VM verifies, and optimality is decided by the exhaustive
solver (pca_amd.Solver), which is also the fallback when the guided search
runs out of budget.
"""
import ctypes as C
import heapq

import numpy as np

from ._lib import PcaGpuError, Task, lib

PER_TEST = 180
H1, H2, EMB, H3 = 512, 256, 16, 256
MAX_T = 32
MAX_DEPTH = 8
REC_WORDS = 16
NDIST = MAX_DEPTH + 1      # remaining steps 0..8
NH = 1 + NDIST             # head outputs: reach logit, remaining-steps logits
DIST_WEIGHT = 0.5          # python/train.py: loss_reach + 0.5 * loss_dist


def _check(rc, what):
    if rc < 0:
        raise PcaGpuError(rc, what)


class DeviceArray:
    """A device buffer of this library's HIP runtime (include/pcavalue.h
    memory helpers) with a numpy dtype and shape; copies are synchronous."""

    def __init__(self, shape, dtype, device=0):
        self.shape = tuple(int(x) for x in (shape if isinstance(shape, (tuple, list)) else (shape,)))
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape)) * self.dtype.itemsize
        self.ptr = lib().pcav_malloc(device, max(self.nbytes, 1))
        self._owner = True
        if not self.ptr:
            raise PcaGpuError(-2, f"device allocation of {self.nbytes} bytes")

    def __del__(self):
        try:
            if self.ptr and self._owner:
                lib().pcav_free(self.ptr)
            self.ptr = None
        except Exception:
            pass

    def view(self, offset_elems, shape):
        """A sub-array sharing this buffer (kept alive by the parent)."""
        v = DeviceArray.__new__(DeviceArray)
        v.shape = tuple(shape)
        v.dtype = self.dtype
        v.nbytes = int(np.prod(shape)) * self.dtype.itemsize
        v.ptr = self.ptr + offset_elems * self.dtype.itemsize
        v._owner = False
        v._parent = self
        return v

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype)
        assert a.nbytes <= self.nbytes
        _check(lib().pcav_copy_to_device(self.ptr, a.ctypes.data, a.nbytes), "copy_to_device")

    def numpy(self):
        out = np.empty(self.shape, dtype=self.dtype)
        _check(lib().pcav_copy_to_host(out.ctypes.data, self.ptr, self.nbytes), "copy_to_host")
        return out

    def zero(self):
        _check(lib().pcav_clear(self.ptr, self.nbytes), "clear")


def state_records(task, progs):
    """Per-test state records (pcagpu_state_records) of kernel prefixes
    `progs` (sequences of instruction words): uint16 [n, 32, 16] and the
    live masks uint8 [n]."""
    n = len(progs)
    ins = np.zeros((max(n, 1), 8), dtype=np.uint16)
    lens = np.zeros(max(n, 1), dtype=np.uint8)
    for i, p in enumerate(progs):
        ins[i, :len(p)] = p
        lens[i] = len(p)
    rec = np.zeros((max(n, 1), MAX_T, REC_WORDS), dtype=np.uint16)
    live = np.zeros(max(n, 1), dtype=np.uint8)
    rc = lib().pcagpu_state_records(C.byref(task), ins.ctypes.data_as(C.POINTER(C.c_uint16)),
                                    lens.ctypes.data_as(C.POINTER(C.c_uint8)), n,
                                    rec.ctypes.data_as(C.POINTER(C.c_uint16)),
                                    live.ctypes.data_as(C.POINTER(C.c_uint8)))
    _check(rc, "state_records")
    return rec[:n], live[:n]


class ValueNet:
    # parameter blocks in one flat FP32 buffer: name, shape (GEMM weights are
    # stored K x N, row-major)
    LAYOUT = [("W1", (PER_TEST, H1)), ("b1", (H1,)), ("W2", (H1, H2)), ("b2", (H2,)),
              ("E", (MAX_DEPTH + 1, EMB)), ("W3", (H2 + EMB, H3)), ("b3", (H3,)), ("W4", (H3, NH)), ("b4", (NH,))]

    def __init__(self, seed=0, device=0):
        self.device = device
        sizes = [int(np.prod(s)) for _, s in self.LAYOUT]
        self.n_params = sum(sizes)
        self.p = DeviceArray(self.n_params, np.float32, device)
        self.g = DeviceArray(self.n_params, np.float32, device)
        self.m = DeviceArray(self.n_params, np.float32, device)
        self.v = DeviceArray(self.n_params, np.float32, device)
        self.p16 = DeviceArray(self.n_params, np.uint16, device)
        for a in (self.g, self.m, self.v):
            a.zero()
        self.views, self.gviews, self.v16 = {}, {}, {}
        off = 0
        rng = np.random.default_rng(seed)
        init = np.zeros(self.n_params, dtype=np.float32)
        for (name, shape), n in zip(self.LAYOUT, sizes):
            self.views[name] = self.p.view(off, shape)
            self.gviews[name] = self.g.view(off, shape)
            self.v16[name] = self.p16.view(off, shape)
            if name.startswith("W"):
                bound = np.sqrt(6.0 / shape[0])          # He-uniform for ReLU layers
                init[off:off + n] = rng.uniform(-bound, bound, n)
            elif name == "E":
                init[off:off + n] = rng.normal(0.0, 0.1, n)
            off += n
        self.p.upload(init)
        self.step_count = 0
        self._sync16()
        self._bufs = {}

    # ------------------------------------------------------------- helpers
    def _sync16(self):
        _check(lib().pcav_to_bf16(self.n_params, self.p.ptr, self.p16.ptr, None), "to_bf16")

    def _buf(self, name, shape, dtype):
        n = int(np.prod(shape))
        t = self._bufs.get(name)
        if t is None or int(np.prod(t.shape)) < n or t.dtype != np.dtype(dtype):
            t = DeviceArray(n, dtype, self.device)
            self._bufs[name] = t
        return t.view(0, shape)

    def _gemm(self, M, N, K, A, lda, at, B, ldb, bt, bias=None, relu=False, C32=None, ldc=0, acc=False,
              C16=None, ldc16=0):
        _check(lib().pcav_gemm(M, N, K, A.ptr, lda, int(at), B.ptr, ldb, int(bt),
                               bias.ptr if bias is not None else None, int(relu),
                               C32.ptr if C32 is not None else None, ldc, int(acc),
                               C16.ptr if C16 is not None else None, ldc16, None), "gemm")

    def parameters(self):
        """FP32 master weights (host copy, flat, LAYOUT order)."""
        return self.p.numpy()

    def gradients(self):
        return self.g.numpy()

    def load_parameters(self, flat):
        self.p.upload(np.asarray(flat, dtype=np.float32))
        self._sync16()

    # ------------------------------------------------------------- forward
    def upload(self, rec, nvalid, depth):
        """Host records -> device inputs of one batch."""
        B = rec.shape[0]
        r = self._buf("rec", (B, MAX_T, REC_WORDS), np.uint16)
        r.upload(rec)
        X = self._buf("X", (B * MAX_T, PER_TEST), np.uint16)
        _check(lib().pcav_features(B, r.ptr, X.ptr, None), "features")
        nv = self._buf("nvalid", (B,), np.uint8)
        nv.upload(np.asarray(nvalid, dtype=np.uint8))
        dp = self._buf("depth", (B,), np.int32)
        dp.upload(np.asarray(depth, dtype=np.int32))
        return X, nv, dp

    def forward(self, X, nv, dp):
        """Head outputs [B, NH] (FP32, device) of a batch: column 0 the reach
        logit, 1.. the remaining-steps logits; keeps the activations for
        backward()."""
        B = nv.shape[0]
        R = B * MAX_T
        w = self.v16
        h1 = self._buf("h1", (R, H1), np.uint16)
        h2 = self._buf("h2", (R, H2), np.uint16)
        z = self._buf("z", (B, H2 + EMB), np.uint16)
        h3 = self._buf("h3", (B, H3), np.uint16)
        y = self._buf("y", (B, NH), np.float32)
        self._gemm(R, H1, PER_TEST, X, PER_TEST, 0, w["W1"], H1, 0, bias=self.views["b1"], relu=True, C16=h1,
                   ldc16=H1)
        self._gemm(R, H2, H1, h1, H1, 0, w["W2"], H2, 0, bias=self.views["b2"], relu=True, C16=h2, ldc16=H2)
        _check(lib().pcav_pool_forward(B, MAX_T, H2, EMB, h2.ptr, nv.ptr, dp.ptr, self.views["E"].ptr, z.ptr,
                                       None), "pool_forward")
        self._gemm(B, H3, H2 + EMB, z, H2 + EMB, 0, w["W3"], H3, 0, bias=self.views["b3"], relu=True, C16=h3,
                   ldc16=H3)
        self._gemm(B, NH, H3, h3, H3, 0, w["W4"], NH, 0, bias=self.views["b4"], C32=y, ldc=NH)
        self._act = (X, nv, dp, h1, h2, z, h3, y, B)
        return y

    # ------------------------------------------------------------ backward
    def backward(self, label, remaining=None, dist_weight=DIST_WEIGHT):
        """Loss of the last forward(): BCE of the reach logit against labels
        (host array [B]) plus dist_weight x cross-entropy of the remaining
        steps (host int array [B], read where the label is positive; None:
        reach only); fills the gradient buffer and returns the loss (host
        float)."""
        X, nv, dp, h1, h2, z, h3, y, B = self._act
        R = B * MAX_T
        w, gv = self.v16, self.gviews
        label = np.asarray(label, dtype=np.float32)
        lab = self._buf("label", (B,), np.float32)
        lab.upload(label)
        n_pos = 0
        rem = self._buf("remaining", (B,), np.int32)
        if remaining is not None and dist_weight:
            rem.upload(np.asarray(remaining, dtype=np.int32))
            n_pos = int((label > 0.5).sum())
        dy = self._buf("dy", (B, NH), np.float32)
        loss = self._buf("loss", (1,), np.float32)
        _check(lib().pcav_reach_dist_loss(B, NH, y.ptr, lab.ptr, rem.ptr, n_pos, float(dist_weight), dy.ptr,
                                          loss.ptr, None), "reach_dist_loss")
        dy16 = self._buf("dy16", (B, NH), np.uint16)
        _check(lib().pcav_to_bf16(B * NH, dy.ptr, dy16.ptr, None), "to_bf16")
        self.g.zero()
        # head: y = h3 W4 + b4
        self._gemm(H3, NH, B, h3, H3, 1, dy16, NH, 0, C32=gv["W4"], ldc=NH)
        _check(lib().pcav_colsum(B, NH, dy.ptr, NH, gv["b4"].ptr, None), "colsum")
        dh3 = self._buf("dh3", (B, H3), np.float32)
        self._gemm(B, H3, NH, dy16, NH, 0, w["W4"], NH, 1, C32=dh3, ldc=H3)
        dh3_16 = self._buf("dh3_16", (B, H3), np.uint16)
        _check(lib().pcav_relu_backward(B * H3, h3.ptr, dh3.ptr, dh3_16.ptr, None), "relu_bwd")
        # h3 = relu(z W3 + b3)
        self._gemm(H2 + EMB, H3, B, z, H2 + EMB, 1, dh3_16, H3, 0, C32=gv["W3"], ldc=H3)
        _check(lib().pcav_colsum(B, H3, dh3.ptr, H3, gv["b3"].ptr, None), "colsum")
        dz = self._buf("dz", (B, H2 + EMB), np.float32)
        self._gemm(B, H2 + EMB, H3, dh3_16, H3, 0, w["W3"], H3, 1, C32=dz, ldc=H2 + EMB)
        # z = [mean_t h2, E[depth]]
        dh2 = self._buf("dh2", (R, H2), np.float32)
        _check(lib().pcav_pool_backward(B, MAX_T, H2, EMB, dz.ptr, nv.ptr, dp.ptr, dh2.ptr, gv["E"].ptr, None),
               "pool_backward")
        dh2_16 = self._buf("dh2_16", (R, H2), np.uint16)
        _check(lib().pcav_relu_backward(R * H2, h2.ptr, dh2.ptr, dh2_16.ptr, None), "relu_bwd")
        # h2 = relu(h1 W2 + b2)
        self._gemm(H1, H2, R, h1, H1, 1, dh2_16, H2, 0, C32=gv["W2"], ldc=H2)
        _check(lib().pcav_colsum(R, H2, dh2.ptr, H2, gv["b2"].ptr, None), "colsum")
        dh1 = self._buf("dh1", (R, H1), np.float32)
        self._gemm(R, H1, H2, dh2_16, H2, 0, w["W2"], H2, 1, C32=dh1, ldc=H1)
        dh1_16 = self._buf("dh1_16", (R, H1), np.uint16)
        _check(lib().pcav_relu_backward(R * H1, h1.ptr, dh1.ptr, dh1_16.ptr, None), "relu_bwd")
        # h1 = relu(X W1 + b1)
        self._gemm(PER_TEST, H1, R, X, PER_TEST, 1, dh1_16, H1, 0, C32=gv["W1"], ldc=H1)
        _check(lib().pcav_colsum(R, H1, dh1.ptr, H1, gv["b1"].ptr, None), "colsum")
        return float(loss.numpy()[0])

    def adam_step(self, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        _check(lib().pcav_adam(self.n_params, self.p.ptr, self.g.ptr, self.m.ptr, self.v.ptr, self.p16.ptr, lr,
                               beta1, beta2, eps, self.step_count, None), "adam")

    # --------------------------------------------------------------- use
    def score(self, task, progs, depth):
        """Reachability logits (host float32) of kernel prefixes of `task`."""
        if not progs:
            return np.zeros(0, dtype=np.float32)
        rec, _ = state_records(task, progs)
        X, nv, dp = self.upload(rec, [task.num_tests] * len(progs), [depth] * len(progs))
        return self.forward(X, nv, dp).numpy()[:, 0].copy()

    def predict_remaining(self, task, progs):
        """The remaining-steps head's choice (0..8) for kernel prefixes (each
        embedded at its own length, as in training)."""
        if not progs:
            return np.zeros(0, dtype=np.int32)
        rec, _ = state_records(task, progs)
        X, nv, dp = self.upload(rec, [task.num_tests] * len(progs), [len(p) for p in progs])
        return self.forward(X, nv, dp).numpy()[:, 1:].argmax(1).astype(np.int32)

    def train(self, rec, nvalid, depth, label, steps, batch=256, lr=1e-3, seed=0, log=None, remaining=None):
        """Adam on minibatches drawn from host arrays (remaining: the
        distance labels, or None for the reach head alone); returns the loss
        curve (every 50 steps and the last)."""
        rng = np.random.default_rng(seed)
        n = rec.shape[0]
        losses = []
        for it in range(steps):
            idx = rng.integers(0, n, size=min(batch, n))
            X, nv, dp = self.upload(rec[idx], nvalid[idx], depth[idx])
            self.forward(X, nv, dp)
            loss = self.backward(label[idx], None if remaining is None else remaining[idx])
            self.adam_step(lr=lr)
            if it % 50 == 0 or it == steps - 1:
                losses.append(loss)
                if log:
                    log(f"step {it} loss {loss:.4f}")
        return losses


# ------------------------------------------------------------------ data
def straight_line_candidates(live, pos, depth):
    """The search grammar at one position without CMP and BR (a straight-line
    program never needs them to produce its output register)."""
    from ._lib import gen_candidates
    return [c for c in gen_candidates(live, pos, depth) if ((c >> 11) & 31) not in (13, 19)]


def _live_after(live, c):
    op, rd = (c >> 11) & 31, (c >> 8) & 7
    if op not in (0, 1, 13, 15, 19) and rd != 0:
        live |= 1 << rd
    return live


def labelled_states(task, depth_opt, solutions, negatives_per_node=4, seed=0):
    """Training states from one exhaustively solved task.  Positives: every
    proper prefix of an optimal program (a solution is reachable in the
    remaining steps), with its distance to a solution, depth_opt minus its
    length (the exhaustive solve proves no shorter completion exists: it
    would be a shorter program).  Negatives: children of those prefixes that
    are not a prefix of any optimal program in the set (distance -1, not
    used).  Returns (progs, depths, labels, remaining)."""
    rng = np.random.default_rng(seed)
    live0 = sum(1 << (i + 1) for i in range(task.num_inputs))
    pos_prefixes = set()
    for prog in solutions:
        for d in range(depth_opt):
            pos_prefixes.add(tuple(prog[:d]))
    progs, depths, labels, remaining = [], [], [], []
    for p in sorted(pos_prefixes):
        progs.append(list(p))
        depths.append(len(p))
        labels.append(1)
        remaining.append(depth_opt - len(p))
        live = live0
        for c in p:
            live = _live_after(live, c)
        cands = straight_line_candidates(live, len(p), depth_opt)
        neg = [c for c in cands if tuple(p) + (c,) not in pos_prefixes and len(p) + 1 < depth_opt]
        for c in rng.permutation(neg)[:negatives_per_node] if neg else []:
            progs.append(list(p) + [int(c)])
            depths.append(len(p) + 1)
            labels.append(0)
            remaining.append(-1)
    return progs, depths, labels, remaining


# ---------------------------------------------------------------- search
def guided_search(task, max_depth, net=None, budget=2000, seed=0, r_exhaustive=0, solver=None, stats=None):
    """Best-first search over straight-line kernel prefixes (no CMP / BR),
    children ordered by the network's logit (net None: by depth, then
    grammar order -- breadth-first, the order without the network).  A
    child whose state has a live register equal to the target on every test
    ends the search; the program is returned with the number of expansions
    (nodes popped within the budget).  Returns (None, expansions) when the
    budget runs out.

    r_exhaustive > 0: a node within r_exhaustive instructions of max_depth
    is not expanded but completed by the GPU's exhaustive leaf completion
    (solver: a pca_amd.Solver of the task), as the reference's loop does
    (python/bootstrap.py:85-113); a completion ends the search, a failed one
    counts as the node's expansion.  stats (a dict) receives the completion
    calls and their time."""
    from . import search_state as ss
    import time as _time
    live0 = sum(1 << (i + 1) for i in range(task.num_inputs))
    tgt = np.array([task.tests[t].out[0] for t in range(task.num_tests)], dtype=np.uint16)
    heap = [(0.0, 0, ())]
    counter = 1
    seen = set()
    expansions = 0
    if stats is not None:
        stats.update(completion_calls=0, completion_s=0.0)
    if r_exhaustive > 0 and solver is None:
        raise ValueError("r_exhaustive needs the task's Solver")
    while heap and expansions < budget:
        _, _, prog = heapq.heappop(heap)
        if len(prog) >= max_depth:
            continue
        if r_exhaustive > 0 and len(prog) >= max_depth - r_exhaustive:
            expansions += 1
            st = ss.after(task, list(prog))
            if st is None:     # a DIV/MOD of the prefix divides by zero on some test
                continue
            t0 = _time.perf_counter()
            found, n, ins = solver.exhaustive_complete(st, max_depth)
            if stats is not None:
                stats["completion_calls"] += 1
                stats["completion_s"] += _time.perf_counter() - t0
            if found:
                end = ss.after(task, list(prog) + ins)
                mask = ss.leaf_check(task, end)
                out_reg = (mask & -mask).bit_length() - 1
                return (list(prog) + ins, out_reg), expansions
            continue
        live = live0
        for c in prog:
            live = _live_after(live, c)
        cands = straight_line_candidates(live, len(prog), max_depth)
        children = [list(prog) + [c] for c in cands]
        rec, lives = state_records(task, children)
        regs = rec[:, :task.num_tests, 3:11]                         # [n, T, 8]
        match = np.all(regs == tgt[None, :, None], axis=1)           # [n, 8]
        faulted = np.any(rec[:, :task.num_tests, 12] != 0, axis=1)   # a test's VM stopped
        expansions += 1
        for i, ch in enumerate(children):
            if faulted[i]:
                continue
            m = match[i] & np.array([(lives[i] >> r) & 1 for r in range(8)], dtype=bool)
            m[0] = False
            if m.any():
                out_reg = int(np.argmax(m))
                return (ch, out_reg), expansions
        # OEP-style dedup of identical states (registers and flags on all tests)
        keys = [rec[i, :task.num_tests, 3:12].tobytes() for i in range(len(children))]
        keep = [i for i, k in enumerate(keys)
                if k not in seen and len(children[i]) < max_depth and not faulted[i]]
        for i in keep:
            seen.add(keys[i])
        if not keep:
            continue
        if net is not None:
            scores = net.score(task, [children[i] for i in keep], len(prog) + 1)
            pri = -scores
        else:
            pri = np.full(len(keep), float(len(prog) + 1))
        for j, i in enumerate(keep):
            heapq.heappush(heap, (float(pri[j]), counter, tuple(children[i])))
            counter += 1
    return None, expansions


def train_on_solved(n_tasks=1200, length=3, seed=101, steps=1500, batch=512, negatives=24, log=None):
    """The network trained on states of exhaustively solved tasks: dependency
    chains of `length` grammar instructions (taskgen.chain_task) solved on
    the GPU (one resident handle, re-tasked), those with an optimum of at
    least 2 labelled by labelled_states (reach, and the distance to a
    solution the solve proves).  Returns (net, info)."""
    import time as _time
    from . import taskgen
    from .solver import Solver
    t0 = _time.perf_counter()
    recs, nvalid, depths, labels, remaining = [], [], [], [], []
    kept = 0
    sv = None
    try:
        for i in range(n_tasks):
            t = taskgen.chain_task(seed, i, length)
            if sv is None:
                sv = Solver(t)
            else:
                sv.retask(t)
            r = sv.solve(length, cap=4096)
            if r.optimal_depth < 2:
                continue
            kept += 1
            d = r.optimal_depth
            kernels = [list(x[0])[t.num_inputs:t.num_inputs + d] for x in r.solutions]
            p, dp, lb, rm = labelled_states(t, d, kernels, negatives_per_node=negatives, seed=len(recs))
            rec, _ = state_records(t, p)
            recs.append(rec)
            nvalid += [t.num_tests] * len(p)
            depths += dp
            labels += lb
            remaining += rm
    finally:
        if sv is not None:
            sv.close()
    data_s = _time.perf_counter() - t0
    net = ValueNet(seed=0)
    t1 = _time.perf_counter()
    losses = net.train(np.concatenate(recs), np.array(nvalid, dtype=np.uint8), np.array(depths, dtype=np.int32),
                       np.array(labels, dtype=np.float32), steps=steps, batch=batch, lr=1e-3,
                       remaining=np.array(remaining, dtype=np.int32), log=log)
    return net, {"tasks": n_tasks, "tasks_kept": kept, "states": len(labels), "positives": int(sum(labels)),
                 "data_s": data_s, "train_s": _time.perf_counter() - t1, "loss_first": float(losses[0]),
                 "loss_last": float(losses[-1])}


def guided_suite(net, n_tasks=100, src_len=(9, 10), max_depth=4, budget=60, seed=29, min_opt=2, scan=8000):
    """Guided against breadth-first best-first search on generated tasks
    whose source programs have src_len instructions (taskgen.random_task),
    with the exhaustive GPU solve to max_depth as the judge and fallback.
    Most such programs reduce to 0 or 1 instructions; the first n_tasks
    whose proven optimum is at least min_opt (of at most `scan` generated)
    are compared: whether each search found a CPU-verified program within
    `budget` expansions, its length and its expansions."""
    import time as _time
    from . import taskgen
    from .solver import Solver
    cases = []
    hist = {}
    scanned = 0
    ex_total = 0.0
    with Solver(taskgen.random_task(seed, 0, src_len)) as sv:
        for i in range(scan):
            if len(cases) >= n_tasks:
                break
            t = taskgen.random_task(seed, i, src_len)
            sv.retask(t)
            t0 = _time.perf_counter()
            r = sv.solve(max_depth, cap=16)
            ex_s = _time.perf_counter() - t0
            ex_total += ex_s
            scanned += 1
            hist[str(r.optimal_depth)] = hist.get(str(r.optimal_depth), 0) + 1
            if r.optimal_depth < min_opt:
                continue
            row = {"task": i, "optimum": r.optimal_depth, "exhaustive_s": ex_s}
            if True:
                for name, m in (("guided", net), ("breadth_first", None)):
                    t0 = _time.perf_counter()
                    res, exp = guided_search(t, max_depth, net=m, budget=budget)
                    ok = res is not None and verify_program(t, res[0], res[1])
                    row[name] = {"found": res is not None, "verified": ok, "length": len(res[0]) if res else None,
                                 "optimal": bool(res is not None and len(res[0]) == r.optimal_depth),
                                 "expansions": exp, "seconds": _time.perf_counter() - t0}
            cases.append(row)
    solved = cases
    out = {"tasks": len(cases), "generated": scanned, "source_length": list(src_len), "max_depth": max_depth,
           "budget": budget, "generator": f"taskgen.random_task(seed {seed}, length {src_len})",
           "optimum_histogram_generated": hist, "optimum_min": min_opt}
    for name in ("guided", "breadth_first"):
        found = [c for c in solved if c[name]["found"]]
        out[name] = {"found": len(found), "verified": sum(1 for c in found if c[name]["verified"]),
                     "optimal": sum(1 for c in solved if c[name]["optimal"]),
                     "mean_expansions_found": (sum(c[name]["expansions"] for c in found) / len(found)) if found else None,
                     "seconds": sum(c[name]["seconds"] for c in solved)}
    both = [c for c in solved if c["guided"]["optimal"] and c["breadth_first"]["optimal"]]
    out["both_optimal"] = len(both)
    out["expansions_on_both"] = {n: sum(c[n]["expansions"] for c in both) for n in ("guided", "breadth_first")}
    out["exhaustive_s"] = ex_total
    out["exhaustive_s_compared"] = sum(c["exhaustive_s"] for c in cases)
    return out


def verify_program(task, prog, out_reg):
    """Runs IN scaffold + prog + OUT out_reg + HLT on the CPU VM over every
    test (pcagpu_task_evaluate); True when all pass."""
    words = [(23 << 11) | ((i + 1) << 8) | task.input_ports[i] for i in range(task.num_inputs)]
    words += list(prog) + [(24 << 11) | (out_reg << 8) | task.output_ports[0], 1 << 11]
    arr = (C.c_uint16 * len(words))(*words)
    return lib().pcagpu_task_evaluate(C.byref(task), arr, len(words)) == task.num_tests
