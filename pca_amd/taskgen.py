"""Seeded synthetic task generators for the benchmark workloads.

The pseudo-random source is xoshiro256** seeded through splitmix64
(Blackman & Vigna), the same published generator the reference's task
generator uses (tools/gen_tasks.c:35-75), so seeds mean the same thing.

Workloads:
  negate_16      one input x, target (-x) mod 2^16; 16 tests: 0, 1, 0x7FFF,
                 0x8000, 0xFFFF and 11 seeded uniform values.
  is_power2_32   one input x, target 1 iff x is a nonzero power of two;
                 32 tests: the 16 powers of two, 0, 0xFFFF and 14 seeded
                 uniform non-powers (no duplicates).
  random_task    target = outputs of a seeded random straight-line program
                 (length 3-8, opcodes uniform over the 25, registers r1-r7,
                 immediates uniform uint16 reduced to the imm8 field), 1 or 2
                 inputs, 8-32 seeded uniform tests.
"""
from ._lib import Task

M64 = (1 << 64) - 1


class Xoshiro256:
    def __init__(self, seed):
        s = []
        x = seed & M64
        for _ in range(4):
            x = (x + 0x9E3779B97F4A7C15) & M64
            z = x
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
            s.append(z ^ (z >> 31))
        self.s = s

    @staticmethod
    def _rotl(x, k):
        return ((x << k) | (x >> (64 - k))) & M64

    def next(self):
        s = self.s
        result = (self._rotl((s[1] * 5) & M64, 7) * 9) & M64
        t = (s[1] << 17) & M64
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = self._rotl(s[3], 45)
        return result

    def u16(self):
        return self.next() & 0xFFFF

    def randint(self, lo, hi):
        return lo + self.next() % (hi - lo + 1)


def negate_16(seed=42):
    rng = Xoshiro256(seed)
    xs = [0, 1, 0x7FFF, 0x8000, 0xFFFF] + [rng.u16() for _ in range(11)]
    return Task.from_io("negate_16", [(x,) for x in xs], [(-x) & 0xFFFF for x in xs], max_cycles=100)


def is_power2(x):
    return 1 if x != 0 and (x & (x - 1)) == 0 else 0


def is_power2_32(seed=42):
    rng = Xoshiro256(seed)
    xs = [1 << i for i in range(16)] + [0, 0xFFFF]
    while len(xs) < 32:
        x = rng.u16()
        if is_power2(x) or x in xs:
            continue
        xs.append(x)
    return Task.from_io("is_power2_32", [(x,) for x in xs], [is_power2(x) for x in xs], max_cycles=100)


# ---- straight-line random programs -----------------------------------------
def _interp(prog, ins, max_cycles=64):
    """Straight-line interpreter of the PCA-16 VM semantics for the
    generator (programs here have no branches or calls)."""
    r = [0] * 8
    flags = 0
    mem = [0] * 256
    ports = [0] * 8
    for i, v in enumerate(ins):
        ports[i] = v
    pc = 0
    for _ in range(max_cycles):
        if pc >= len(prog):
            return None
        w = prog[pc]
        pc += 1
        op, rd, rs, rt, imm = (w >> 11) & 31, (w >> 8) & 7, (w >> 5) & 7, (w >> 2) & 7, w & 0xFF
        a, b = r[rs], r[rt]
        if op == 1:
            return ports[0]
        if op in (2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12):
            if op in (5, 6) and b == 0:
                return None
            sa = a - 0x10000 if a & 0x8000 else a
            sb = b - 0x10000 if b & 0x8000 else b
            if op == 2: v = a + b
            elif op == 3: v = a - b
            elif op == 4: v = a * b
            elif op == 5: v = int(sa / sb)
            elif op == 6: v = sa - int(sa / sb) * sb
            elif op == 7: v = a & b
            elif op == 8: v = a | b
            elif op == 9: v = a ^ b
            elif op == 10: v = a << (b & 15)
            elif op == 11: v = a >> (b & 15)
            else: v = sa >> (b & 15)
            r[rd] = v & 0xFFFF
        elif op == 14:
            r[rd] = mem[r[rs] & 0xFF]
        elif op == 15:
            mem[r[rd] & 0xFF] = r[rs]
        elif op == 16:
            r[rd] = imm
        elif op == 17:
            r[rd] = (r[rd] & 0xFF) | (imm << 8)
        elif op == 18:
            r[rd] = (r[rd] + (imm - 256 if imm & 0x80 else imm)) & 0xFFFF
        elif op == 23:
            r[rd] = ports[imm & 7]
        elif op == 24:
            ports[imm & 7] = r[rd]
        elif op in (0, 13, 19, 20, 21, 22):
            pass   # NOP / CMP / control flow: no register effect on this line
        else:
            return None
        r[0] = 0
    return None


def random_task(seed, index, length=(3, 8)):
    """Task #index of the seeded synthetic suite (see module docstring).

    Control-flow opcodes (BR, JMP, CALL, RET) are drawn like the others but
    executed as no-ops on the generating straight line, and an immediate
    drawn uniform over uint16 is reduced to its low byte, the width of the
    imm8 field.  The generating program's OUT r, 0 of a random register
    and HLT end it."""
    rng = Xoshiro256((seed * 1000003 + index) & M64)
    n_in = rng.randint(1, 2)
    length = rng.randint(length[0], length[1])
    prog = [(23 << 11) | ((i + 1) << 8) | i for i in range(n_in)]
    for _ in range(length):
        op = rng.randint(0, 24)
        rd, rs, rt = rng.randint(1, 7), rng.randint(1, 7), rng.randint(1, 7)
        imm = rng.u16() & 0xFF
        if op in (1, 23, 24):     # keep HLT/IN/OUT for the scaffold and tail
            op = 0
        if op in (16, 17, 18, 19):
            w = (op << 11) | (rd << 8) | imm
        else:
            w = (op << 11) | (rd << 8) | (rs << 5) | (rt << 2)
        prog.append(w)
    out_reg = rng.randint(1, 7)
    prog += [(24 << 11) | (out_reg << 8), 1 << 11]
    n_tests = rng.randint(8, 32)
    ins, outs = [], []
    tries = 0
    while len(ins) < n_tests and tries < 4 * n_tests:
        tries += 1
        x = tuple(rng.u16() for _ in range(n_in))
        y = _interp(prog, x)
        if y is None:
            continue
        ins.append(x)
        outs.append(y)
    if not ins:
        ins, outs = [tuple(0 for _ in range(n_in))], [0]
    return Task.from_io(f"syn_{seed}_{index}", ins, outs, max_cycles=100)


def _shallow_draw(rng, max_len, gen_candidates):
    """One program draw of shallow_task and its tests; None when 256 inputs
    in a row divide by zero (a divisor that is constant zero), in which case
    the caller draws another program from the same stream."""
    n_in = rng.randint(1, 2)
    live = sum(1 << (i + 1) for i in range(n_in))
    length = rng.randint(1, max_len)
    prog = [(23 << 11) | ((i + 1) << 8) | i for i in range(n_in)]
    for pos in range(length):
        cands = [c for c in gen_candidates(live, pos, length)
                 if ((c >> 11) & 31) not in (13, 19)]          # no CMP / BR
        c = cands[rng.next() % len(cands)]
        prog.append(c)
        live |= 1 << ((c >> 8) & 7)
    regs = [r for r in range(1, 8) if live >> r & 1]
    out_reg = regs[rng.next() % len(regs)]
    prog += [(24 << 11) | (out_reg << 8), 1 << 11]
    n_tests = rng.randint(8, 32)
    ins, outs = [], []
    misses = 0
    while len(ins) < n_tests:
        x = tuple(rng.u16() for _ in range(n_in))
        y = _interp(prog, x)
        if y is None:          # DIV/MOD by zero on this input: draw another
            misses += 1
            if misses == 256:
                return None
            continue
        misses = 0
        ins.append(x)
        outs.append(y)
    return ins, outs


def chain_task(seed, index, length=3):
    """A task whose generating program is a dependency chain of `length`
    search-grammar instructions (ALU, LDI, ADDI; no CMP / BR): each reads the
    register the previous one wrote and the last one's register is the
    output, so the optimum is rarely much shorter than the source (the
    training and evaluation tasks of the value network).  1 or 2 inputs,
    8-32 seeded uniform tests."""
    from ._lib import gen_candidates
    rng = Xoshiro256((seed * 6151 + index * 92821 + 5) & M64)
    while True:
        n_in = rng.randint(1, 2)
        live = sum(1 << (i + 1) for i in range(n_in))
        last = rng.randint(1, n_in)
        prog = [(23 << 11) | ((i + 1) << 8) | i for i in range(n_in)]
        for pos in range(length):
            cands = []
            for c in gen_candidates(live, pos, length):
                op, rd, rs, rt = (c >> 11) & 31, (c >> 8) & 7, (c >> 5) & 7, (c >> 2) & 7
                if op in (13, 19) or rd == 0:
                    continue
                reads_last = (op == 18 and rd == last) or (op not in (16, 18) and last in (rs, rt))
                if reads_last:
                    cands.append(c)
            c = cands[rng.next() % len(cands)]
            prog.append(c)
            last = (c >> 8) & 7
            live |= 1 << last
        prog += [(24 << 11) | (last << 8), 1 << 11]
        n_tests = rng.randint(8, 32)
        ins, outs, misses = [], [], 0
        while len(ins) < n_tests and misses < 256:
            x = tuple(rng.u16() for _ in range(n_in))
            y = _interp(prog, x)
            if y is None:
                misses += 1
                continue
            misses = 0
            ins.append(x)
            outs.append(y)
        if len(ins) == n_tests:
            return Task.from_io(f"chain_{seed}_{index}", ins, outs, max_cycles=100)


def shallow_task(seed, index, max_len=3):
    """A serving request: the outputs of a seeded random program of 1..max_len
    search-grammar instructions (ALU, LDI, ADDI over live registers) on 1 or 2
    inputs and 8-32 seeded uniform tests, so the optimum is at most max_len
    kernel instructions deep."""
    from ._lib import gen_candidates
    rng = Xoshiro256((seed * 7919 + index * 104729 + 17) & M64)
    while True:
        drawn = _shallow_draw(rng, max_len, gen_candidates)
        if drawn is not None:
            ins, outs = drawn
            break
    return Task.from_io(f"srv_{seed}_{index}", ins, outs, max_cycles=100)
