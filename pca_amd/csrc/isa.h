// isa.h -- PCA-16 instruction set and the straight-line kernel step, shared
// by the host prefix producer and the gfx950 kernels (one semantics).
//
// Encoding and opcodes follow src/pca.h:32-102; flag rules, r0 clamp and the
// signed DIV/MOD rules follow src/vm.c:28-205.  The step below is the VM
// restricted to what a kernel position of the search grammar can hold (ALU,
// CMP, forward BR, LDI, ADDI) plus the two ways a test can leave the
// straight line: a taken forward BR (execution resumes at a later kernel
// position) and a DIV/MOD fault (the VM stops; state frozen).  That is
// exactly the state the reference's OEP replay observes
// (tools/enumerate.c:435-452: scaffold + kernel + HLT, run 100 cycles,
// status ignored).
#pragma once
#include <stdint.h>

#if defined(__HIP__)
#include <hip/hip_runtime.h>
#define PCA_HD __host__ __device__ __forceinline__
#else
#define PCA_HD static inline
#endif

namespace pca {

enum : uint32_t {
    OP_NOP = 0x00, OP_HLT, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MOD, OP_AND, OP_OR,
    OP_XOR, OP_SHL, OP_SHR, OP_ASR, OP_CMP, OP_LD, OP_ST, OP_LDI, OP_LUI, OP_ADDI,
    OP_BR, OP_JMP, OP_CALL, OP_RET, OP_IN, OP_OUT, OP_COUNT
};

constexpr uint32_t FLAG_Z = 1, FLAG_N = 2, FLAG_C = 4, FLAG_V = 8;

PCA_HD uint32_t insn_op(uint32_t i) { return (i >> 11) & 31u; }
PCA_HD uint32_t insn_rd(uint32_t i) { return (i >> 8) & 7u; }
PCA_HD uint32_t insn_rs(uint32_t i) { return (i >> 5) & 7u; }
PCA_HD uint32_t insn_rt(uint32_t i) { return (i >> 2) & 7u; }
PCA_HD uint32_t insn_imm(uint32_t i) { return i & 0xFFu; }
PCA_HD int32_t insn_simm(uint32_t i) { return (int32_t)(int8_t)(uint8_t)(i & 0xFFu); }

PCA_HD uint16_t enc_r(uint32_t op, uint32_t rd, uint32_t rs, uint32_t rt)
{
    return (uint16_t)(((op & 31u) << 11) | ((rd & 7u) << 8) | ((rs & 7u) << 5) | ((rt & 7u) << 2));
}
PCA_HD uint16_t enc_i(uint32_t op, uint32_t rd, uint32_t imm)
{
    return (uint16_t)(((op & 31u) << 11) | ((rd & 7u) << 8) | (imm & 0xFFu));
}

PCA_HD uint32_t flags_zn(uint32_t r)
{
    return ((r & 0xFFFFu) == 0 ? FLAG_Z : 0u) | ((r & 0x8000u) ? FLAG_N : 0u);
}
PCA_HD uint32_t flags_add(uint32_t a, uint32_t b)
{
    uint32_t full = a + b, r = full & 0xFFFFu;
    return flags_zn(r) | (full > 0xFFFFu ? FLAG_C : 0u) |
           ((~(a ^ b) & (a ^ r) & 0x8000u) ? FLAG_V : 0u);
}
PCA_HD uint32_t flags_sub(uint32_t a, uint32_t b)
{
    uint32_t r = (a - b) & 0xFFFFu;
    return flags_zn(r) | (a >= b ? FLAG_C : 0u) | (((a ^ b) & (a ^ r) & 0x8000u) ? FLAG_V : 0u);
}

// The quotient x / y of C's int division (truncated toward zero) for x, y
// in [-32768, 32767], y != 0.  On the device it comes from a float
// reciprocal instead of the integer-division sequence: x and y are exact in
// float; v_rcp_f32's estimate of 1/y (within one ulp), raised by the factor
// 1 + 2^-20, times x lies above |x / y| in magnitude (so an integer quotient
// is not truncated to the integer below it) by less than 2^-4 / |y|, which
// is less than the distance from any non-integer quotient (a multiple of
// 1 / |y|) to the next integer, so truncation gives the exact quotient.
// Checked over all 2^32 operand pairs against the CPU VM on the device, and
// on the CPU with the reciprocal rounded to nearest and one ulp either side.
PCA_HD int32_t sdiv16_q(int32_t x, int32_t y)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const float r = __builtin_amdgcn_rcpf((float)y) * (1.0f + 0x1p-20f);
    return (int32_t)((float)x * r);
#else
    return x / y;
#endif
}

// C semantics of (int16_t)a / (int16_t)b and % after int promotion
// (src/vm.c:94-103): truncation toward zero, -32768 / -1 = 32768 -> 0x8000,
// -32768 % -1 = 0.  b != 0 is the caller's job.
PCA_HD uint32_t sdiv16(uint32_t a, uint32_t b)
{
    const int32_t x = (int16_t)(uint16_t)a, y = (int16_t)(uint16_t)b;
    return (uint32_t)sdiv16_q(x, y) & 0xFFFFu;
}
PCA_HD uint32_t smod16(uint32_t a, uint32_t b)
{
    const int32_t x = (int16_t)(uint16_t)a, y = (int16_t)(uint16_t)b;
    return (uint32_t)(x - sdiv16_q(x, y) * y) & 0xFFFFu;
}

// BR condition codes (src/pca.h:62-70, src/vm.c:159-172).  Code 7 is never
// taken, as on the CPU VM.
PCA_HD bool cond_taken(uint32_t cond, uint32_t f)
{
    bool z = f & FLAG_Z, n = f & FLAG_N, c = f & FLAG_C, v = f & FLAG_V;
    switch (cond) {
    case 0: return true;
    case 1: return z;
    case 2: return !z;
    case 3: return n != v;
    case 4: return n == v;
    case 5: return !c;
    case 6: return c;
    default: return false;
    }
}

// Value written by a register-writing grammar instruction and its flags.
// For LDI flags are unchanged (returned as `f`).  *fault = DIV/MOD by zero.
PCA_HD uint32_t alu_eval(uint32_t op, uint32_t a, uint32_t b, uint32_t imm, uint32_t f,
                         uint32_t *fl_out, bool *fault)
{
    uint32_t r = 0;
    *fault = false;
    switch (op) {
    case OP_ADD: r = (a + b) & 0xFFFFu; f = flags_add(a, b); break;
    case OP_SUB: r = (a - b) & 0xFFFFu; f = flags_sub(a, b); break;
    case OP_MUL: {
        uint32_t p = a * b;
        r = p & 0xFFFFu;
        f = flags_zn(r) | (p > 0xFFFFu ? FLAG_C : 0u);
        break;
    }
    case OP_DIV:
        if (b == 0) { *fault = true; break; }
        r = sdiv16(a, b); f = flags_zn(r); break;
    case OP_MOD:
        if (b == 0) { *fault = true; break; }
        r = smod16(a, b); f = flags_zn(r); break;
    case OP_AND: r = a & b; f = flags_zn(r); break;
    case OP_OR:  r = a | b; f = flags_zn(r); break;
    case OP_XOR: r = a ^ b; f = flags_zn(r); break;
    case OP_SHL: r = (a << (b & 15u)) & 0xFFFFu; f = flags_zn(r); break;
    case OP_SHR: r = a >> (b & 15u); f = flags_zn(r); break;
    case OP_ASR: r = (uint32_t)((int32_t)(int16_t)(uint16_t)a >> (b & 15u)) & 0xFFFFu; f = flags_zn(r); break;
    case OP_LDI: r = imm; break;
    case OP_ADDI: {
        uint32_t ext = (uint32_t)(int32_t)(int8_t)(uint8_t)imm & 0xFFFFu;
        r = (a + ext) & 0xFFFFu;
        f = flags_add(a, ext);
        break;
    }
    default: break;
    }
    *fl_out = f;
    return r;
}

// One test's state on the straight line of a kernel.  `resume` is the kernel
// position where a taken forward BR lands; positions below it are skipped.
struct alignas(4) TestState {
    uint16_t r[8];
    uint8_t flags;
    uint8_t resume;
    uint8_t frozen;   // DIV/MOD by zero happened: the VM stopped
    uint8_t exec;     // instructions executed so far, scaffold included
};

// Execute kernel instruction `insn` at kernel position `pos`.
PCA_HD void step_test(TestState &s, uint32_t insn, uint32_t pos)
{
    if (s.frozen || s.resume > pos) return;
    s.exec++;
    uint32_t op = insn_op(insn), rd = insn_rd(insn), rs = insn_rs(insn), rt = insn_rt(insn);
    uint32_t imm = insn_imm(insn);
    if (op == OP_BR) {
        if (cond_taken(rd, s.flags)) s.resume = (uint8_t)(pos + 1 + (uint32_t)insn_simm(insn));
        return;
    }
    if (op == OP_CMP) {
        s.flags = (uint8_t)flags_sub(s.r[rd], s.r[rs]);
        return;
    }
    uint32_t a = (op == OP_ADDI) ? s.r[rd] : s.r[rs];
    uint32_t b = s.r[rt];
    uint32_t f;
    bool fault;
    uint32_t v = alu_eval(op, a, b, imm, s.flags, &f, &fault);
    if (fault) { s.frozen = 1; return; }
    s.r[rd] = (uint16_t)v;
    s.flags = (uint8_t)f;
    s.r[0] = 0;
}

// The same state as five 32-bit words in the TestState byte layout: w[0..3]
// hold r0..r7 (r[2i] in the low half of w[i]), w[4] = flags | resume << 8 |
// frozen << 16 | exec << 24.  Kernels keep it in registers and step it with
// the instruction (and so its register fields) uniform across the wave;
// step_packed is step_test on this form.
struct PackedState {
    uint32_t w[5];
};

PCA_HD uint32_t ps_reg(const PackedState &s, uint32_t k)
{
    const uint32_t d = k >> 1;
    const uint32_t v = d == 0 ? s.w[0] : (d == 1 ? s.w[1] : (d == 2 ? s.w[2] : s.w[3]));
    return (k & 1u) ? v >> 16 : v & 0xFFFFu;
}

// (written as a chain on the word index, not a loop over the words: a loop
// makes the compiler keep the state in a memory array to index it)
PCA_HD void ps_set_reg(PackedState &s, uint32_t k, uint32_t v)
{
    const uint32_t sh = (k & 1u) * 16u, keep = ~(0xFFFFu << sh), d = k >> 1;
    const uint32_t nv = (v & 0xFFFFu) << sh;
    if (d == 0)
        s.w[0] = (s.w[0] & keep) | nv;
    else if (d == 1)
        s.w[1] = (s.w[1] & keep) | nv;
    else if (d == 2)
        s.w[2] = (s.w[2] & keep) | nv;
    else
        s.w[3] = (s.w[3] & keep) | nv;
}

// Whether a test runs the instruction at pos: not frozen, no pending branch
// past pos.
PCA_HD bool ps_runs(uint32_t w4, uint32_t pos) { return (w4 & 0xFF0000u) == 0u && ((w4 >> 8) & 0xFFu) <= pos; }

// FLAGS = false: the flags byte is left as it was (for a state whose flags
// nothing reads: the last kernel position holds no BR, and its value-writing
// instructions do not read flags); registers, resume, frozen and exec are
// stepped exactly.  CHECK = false: the caller has checked ps_runs.
template <bool FLAGS = true, bool CHECK = true>
PCA_HD void step_packed(PackedState &s, uint32_t insn, uint32_t pos)
{
    const uint32_t w4 = s.w[4];
    const uint32_t flags = w4 & 0xFFu;
    if (CHECK && !ps_runs(w4, pos)) return;
    uint32_t nw4 = w4 + (1u << 24);   // exec++ (a byte: wraps like the uint8_t field)
    const uint32_t op = insn_op(insn), rd = insn_rd(insn), rs = insn_rs(insn), rt = insn_rt(insn);
    if (op == OP_BR) {
        if (cond_taken(rd, flags))
            nw4 = (nw4 & ~0xFF00u) | (((pos + 1u + (uint32_t)insn_simm(insn)) & 0xFFu) << 8);
        s.w[4] = nw4;
        return;
    }
    if (op == OP_CMP) {
        s.w[4] = FLAGS ? (nw4 & ~0xFFu) | flags_sub(ps_reg(s, rd), ps_reg(s, rs)) : nw4;
        return;
    }
    const uint32_t a = ps_reg(s, op == OP_ADDI ? rd : rs), b = ps_reg(s, rt);
    uint32_t f;
    bool fault;
    const uint32_t v = alu_eval(op, a, b, insn_imm(insn), flags, &f, &fault);
    if (fault) {
        s.w[4] = nw4 | (1u << 16);
        return;
    }
    ps_set_reg(s, rd, v);
    s.w[0] &= 0xFFFF0000u;   // r0 stays 0
    s.w[4] = FLAGS ? (nw4 & ~0xFFu) | (f & 0xFFu) : nw4;
}

// Live mask after an instruction (tools/enumerate.c:427-433).
PCA_HD uint32_t live_after(uint32_t live, uint32_t insn)
{
    uint32_t op = insn_op(insn), rd = insn_rd(insn);
    if (op != OP_CMP && op != OP_BR && op != OP_NOP && op != OP_HLT && op != OP_ST && rd != 0)
        live |= 1u << rd;
    return live;
}

}  // namespace pca
