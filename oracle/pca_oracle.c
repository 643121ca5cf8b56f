/*
 * pca_oracle.c -- CPU restatement of the PCA-16 reference solver.
 * TEST INFRASTRUCTURE ONLY (see pca_oracle.h).  Each function names the
 * reference lines it follows.
 */
#define _POSIX_C_SOURCE 199309L
#include "pca_oracle.h"

#include <ctype.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* opcodes, conditions, flags: src/pca.h:32-82 */
enum {
    O_NOP = 0x00, O_HLT, O_ADD, O_SUB, O_MUL, O_DIV, O_MOD, O_AND, O_OR,
    O_XOR, O_SHL, O_SHR, O_ASR, O_CMP, O_LD, O_ST, O_LDI, O_LUI, O_ADDI,
    O_BR, O_JMP, O_CALL, O_RET, O_IN, O_OUT, O_COUNT
};
#define F_Z 0x01
#define F_N 0x02
#define F_C 0x04
#define F_V 0x08

/* encoding: src/pca.h:91-102 */
#define I_OP(i)   (((i) >> 11) & 0x1F)
#define I_RD(i)   (((i) >> 8) & 0x07)
#define I_RS(i)   (((i) >> 5) & 0x07)
#define I_RT(i)   (((i) >> 2) & 0x07)
#define I_IMM(i)  ((i) & 0xFF)
#define ENC_R(op, rd, rs, rt) \
    ((uint16_t)(((op)&0x1F) << 11 | ((rd)&7) << 8 | ((rs)&7) << 5 | ((rt)&7) << 2))
#define ENC_I(op, rd, imm) \
    ((uint16_t)(((op)&0x1F) << 11 | ((rd)&7) << 8 | ((imm)&0xFF)))

/* ------------------------------------------------------------------ VM */
/* The reference compiles the VM as its own unit (src/vm.c), so the solver
 * calls pca_init/pca_load/pca_step/pca_run out of line.  NOINLINE keeps the
 * same call structure here so the CPU-baseline rate is comparable. */
#define NOINLINE __attribute__((noinline))

/* src/vm.c:14-24 */
NOINLINE void or_vm_init(or_vm_t *vm) { memset(vm, 0, sizeof(*vm)); }

NOINLINE void or_vm_load(or_vm_t *vm, const uint16_t *prog, size_t len)
{
    if (len > OR_PROG_SIZE) len = OR_PROG_SIZE;
    memcpy(vm->prog, prog, len * sizeof(uint16_t));
}

/* flag helpers: src/vm.c:28-54 */
static uint8_t fl_zn(uint16_t r)
{
    return (uint8_t)((r == 0 ? F_Z : 0) | ((r & 0x8000) ? F_N : 0));
}
static uint8_t fl_add(uint16_t a, uint16_t b)
{
    uint32_t full = (uint32_t)a + (uint32_t)b;
    uint16_t r = (uint16_t)full;
    uint8_t f = fl_zn(r);
    if (full > 0xFFFF) f |= F_C;
    if (~(a ^ b) & (a ^ r) & 0x8000) f |= F_V;
    return f;
}
static uint8_t fl_sub(uint16_t a, uint16_t b)
{
    uint16_t r = (uint16_t)(a - b);
    uint8_t f = fl_zn(r);
    if (a >= b) f |= F_C;
    if ((a ^ b) & (a ^ r) & 0x8000) f |= F_V;
    return f;
}

uint16_t or_alu(int op, uint16_t a, uint16_t b, uint8_t flags_in,
                uint8_t *flags_out, int *fault)
{
    uint16_t r = 0;
    uint8_t f = flags_in;
    *fault = 0;
    switch (op) {
    case O_ADD: f = fl_add(a, b); r = (uint16_t)(a + b); break;
    case O_SUB: f = fl_sub(a, b); r = (uint16_t)(a - b); break;
    case O_MUL: {
        uint32_t p = (uint32_t)a * (uint32_t)b;
        r = (uint16_t)p;
        f = fl_zn(r);
        if (p > 0xFFFF) f |= F_C;
        break;
    }
    case O_DIV:
        if (b == 0) { *fault = 1; break; }
        r = (uint16_t)((int)(int16_t)a / (int)(int16_t)b);
        f = fl_zn(r);
        break;
    case O_MOD:
        if (b == 0) { *fault = 1; break; }
        r = (uint16_t)((int)(int16_t)a % (int)(int16_t)b);
        f = fl_zn(r);
        break;
    case O_AND: r = a & b; f = fl_zn(r); break;
    case O_OR:  r = a | b; f = fl_zn(r); break;
    case O_XOR: r = a ^ b; f = fl_zn(r); break;
    case O_SHL: r = (uint16_t)(a << (b & 0xF)); f = fl_zn(r); break;
    case O_SHR: r = (uint16_t)(a >> (b & 0xF)); f = fl_zn(r); break;
    case O_ASR: r = (uint16_t)((int)(int16_t)a >> (b & 0xF)); f = fl_zn(r); break;
    case O_CMP: f = fl_sub(a, b); break;
    default: break;
    }
    *flags_out = f;
    return r;
}

/* src/vm.c:58-205: same statement shape as the reference step (flag helpers
 * writing the flags byte, callback checks on IN/OUT) so that the port's
 * node rate tracks the reference binary's. */
NOINLINE void or_vm_step(or_vm_t *vm)
{
    if (vm->status != OR_RUNNING) return;
    uint16_t insn = vm->prog[vm->pc];
    uint8_t op = I_OP(insn), rd = I_RD(insn), rs = I_RS(insn), rt = I_RT(insn);
    uint8_t imm = I_IMM(insn);
    int8_t simm = (int8_t)imm;
    vm->pc++;
    switch (op) {
    case O_NOP: break;
    case O_HLT: vm->status = OR_HALTED; break;
    case O_ADD:
        vm->flags = fl_add(vm->r[rs], vm->r[rt]);
        vm->r[rd] = (uint16_t)(vm->r[rs] + vm->r[rt]);
        break;
    case O_SUB:
        vm->flags = fl_sub(vm->r[rs], vm->r[rt]);
        vm->r[rd] = (uint16_t)(vm->r[rs] - vm->r[rt]);
        break;
    case O_MUL: {
        uint32_t prod = (uint32_t)vm->r[rs] * (uint32_t)vm->r[rt];
        vm->r[rd] = (uint16_t)prod;
        vm->flags = fl_zn(vm->r[rd]);
        if (prod > 0xFFFF) vm->flags |= F_C;
        break;
    }
    case O_DIV:
        if (vm->r[rt] == 0) { vm->status = OR_FAULT; break; }
        vm->r[rd] = (uint16_t)((int)(int16_t)vm->r[rs] / (int)(int16_t)vm->r[rt]);
        vm->flags = fl_zn(vm->r[rd]);
        break;
    case O_MOD:
        if (vm->r[rt] == 0) { vm->status = OR_FAULT; break; }
        vm->r[rd] = (uint16_t)((int)(int16_t)vm->r[rs] % (int)(int16_t)vm->r[rt]);
        vm->flags = fl_zn(vm->r[rd]);
        break;
    case O_AND: vm->r[rd] = vm->r[rs] & vm->r[rt]; vm->flags = fl_zn(vm->r[rd]); break;
    case O_OR:  vm->r[rd] = vm->r[rs] | vm->r[rt]; vm->flags = fl_zn(vm->r[rd]); break;
    case O_XOR: vm->r[rd] = vm->r[rs] ^ vm->r[rt]; vm->flags = fl_zn(vm->r[rd]); break;
    case O_SHL: vm->r[rd] = (uint16_t)(vm->r[rs] << (vm->r[rt] & 0xF)); vm->flags = fl_zn(vm->r[rd]); break;
    case O_SHR: vm->r[rd] = (uint16_t)(vm->r[rs] >> (vm->r[rt] & 0xF)); vm->flags = fl_zn(vm->r[rd]); break;
    case O_ASR:
        vm->r[rd] = (uint16_t)((int)(int16_t)vm->r[rs] >> (vm->r[rt] & 0xF));
        vm->flags = fl_zn(vm->r[rd]);
        break;
    case O_CMP: vm->flags = fl_sub(vm->r[rd], vm->r[rs]); break;
    case O_LD: vm->r[rd] = vm->mem[vm->r[rs] & 0xFF]; break;
    case O_ST: vm->mem[vm->r[rd] & 0xFF] = vm->r[rs]; break;
    case O_LDI: vm->r[rd] = imm; break;
    case O_LUI: vm->r[rd] = (uint16_t)((vm->r[rd] & 0x00FF) | ((uint16_t)imm << 8)); break;
    case O_ADDI: {
        uint16_t ext = (uint16_t)(int16_t)simm;
        vm->flags = fl_add(vm->r[rd], ext);
        vm->r[rd] = (uint16_t)(vm->r[rd] + ext);
        break;
    }
    case O_BR: {
        int take = 0;
        uint8_t f = vm->flags;
        switch (rd) {
        case 0: take = 1; break;
        case 1: take = (f & F_Z) != 0; break;
        case 2: take = (f & F_Z) == 0; break;
        case 3: take = ((f & F_N) != 0) != ((f & F_V) != 0); break;
        case 4: take = ((f & F_N) != 0) == ((f & F_V) != 0); break;
        case 5: take = (f & F_C) == 0; break;
        case 6: take = (f & F_C) != 0; break;
        default: break; /* cond 7: not taken on the CPU VM */
        }
        if (take) vm->pc = (uint8_t)((int)vm->pc + simm);
        break;
    }
    case O_JMP: vm->pc = (uint8_t)vm->r[rd]; break;
    case O_CALL:
        if (vm->sp >= 16) { vm->status = OR_FAULT; break; }
        vm->stack[vm->sp++] = vm->pc;
        vm->pc = (uint8_t)vm->r[rd];
        break;
    case O_RET:
        if (vm->sp == 0) { vm->status = OR_FAULT; break; }
        vm->pc = (uint8_t)vm->stack[--vm->sp];
        break;
    case O_IN:
        if (vm->on_read) vm->on_read(vm, imm & 7);
        vm->r[rd] = vm->ports[imm & 7];
        break;
    case O_OUT:
        vm->ports[imm & 7] = vm->r[rd];
        if (vm->on_write) vm->on_write(vm, imm & 7, vm->r[rd]);
        break;
    default: vm->status = OR_FAULT; break;
    }
    vm->r[0] = 0;
    vm->cycles++;
}

/* src/vm.c:209-217 */
NOINLINE int or_vm_run(or_vm_t *vm, uint64_t max_cycles)
{
    while (vm->status == OR_RUNNING && vm->cycles < max_cycles) or_vm_step(vm);
    if (vm->status == OR_RUNNING && vm->cycles >= max_cycles) vm->status = OR_TIMEOUT;
    return vm->status;
}

/* ---------------------------------------------------------- task loader */
/* tasks/spec.c:64-127: a strstr-based reader for the task schema only. */
static const char *sk_ws(const char *p)
{
    while (*p && isspace((unsigned char)*p)) p++;
    return p;
}
static const char *fkey(const char *json, const char *key)
{
    char pat[128];
    snprintf(pat, sizeof(pat), "\"%s\"", key);
    const char *p = strstr(json, pat);
    if (!p) return NULL;
    p = sk_ws(p + strlen(pat));
    if (*p == ':') p++;
    return sk_ws(p);
}
static const char *pstr(const char *p, char *out, int maxlen)
{
    p = sk_ws(p);
    if (*p != '"') return NULL;
    p++;
    int i = 0;
    while (*p && *p != '"' && i < maxlen - 1) out[i++] = *p++;
    out[i] = 0;
    if (*p == '"') p++;
    return p;
}
static const char *pint(const char *p, long *out)
{
    char *end;
    p = sk_ws(p);
    *out = strtol(p, &end, 10);
    return end;
}
static const char *parr(const char *p, long *arr, int maxlen, int *count)
{
    p = sk_ws(p);
    if (*p != '[') return NULL;
    p++;
    *count = 0;
    while (*count < maxlen) {
        p = sk_ws(p);
        if (*p == ']') return p + 1;
        if (*p == ',') p++;
        p = pint(p, &arr[*count]);
        (*count)++;
    }
    while (*p && *p != ']') p++;
    if (*p == ']') p++;
    return p;
}

/* tasks/spec.c:131-235 (the parse part) */
int or_task_parse(const char *json, or_task_t *t)
{
    memset(t, 0, sizeof(*t));
    const char *p = fkey(json, "name");
    if (p) pstr(p, t->name, sizeof(t->name));
    p = fkey(json, "max_cycles");
    if (p) {
        long v;
        pint(p, &v);
        t->max_cycles = (uint64_t)v;
    } else {
        t->max_cycles = 1000;
    }
    p = fkey(json, "input_ports");
    if (p) {
        long ports[OR_MAX_PORTS];
        int cnt = 0;
        parr(p, ports, OR_MAX_PORTS, &cnt);
        t->num_inputs = cnt;
        for (int i = 0; i < cnt; i++) t->input_ports[i] = (uint8_t)ports[i];
    }
    p = fkey(json, "output_ports");
    if (p) {
        long ports[OR_MAX_PORTS];
        int cnt = 0;
        parr(p, ports, OR_MAX_PORTS, &cnt);
        t->num_outputs = cnt;
        for (int i = 0; i < cnt; i++) t->output_ports[i] = (uint8_t)ports[i];
    }
    p = fkey(json, "tests");
    if (p) {
        p = sk_ws(p);
        if (*p == '[') p++;
        t->num_tests = 0;
        while (t->num_tests < OR_MAX_TESTS) {
            p = sk_ws(p);
            if (*p == ']' || *p == 0) break;
            if (*p == ',') p++;
            p = sk_ws(p);
            if (*p != '{') break;
            const char *obj = p;
            int depth = 1;
            p++;
            while (*p && depth > 0) {
                if (*p == '{') depth++;
                else if (*p == '}') depth--;
                p++;
            }
            int ti = t->num_tests;
            const char *ip = fkey(obj, "in");
            if (ip) {
                long v[OR_MAX_PORTS];
                int cnt = 0;
                parr(ip, v, OR_MAX_PORTS, &cnt);
                for (int i = 0; i < cnt; i++) t->tests[ti].in[i] = (uint16_t)(v[i] & 0xFFFF);
            }
            const char *op = fkey(obj, "out");
            if (op) {
                long v[OR_MAX_PORTS];
                int cnt = 0;
                parr(op, v, OR_MAX_PORTS, &cnt);
                for (int i = 0; i < cnt; i++) t->tests[ti].out[i] = (uint16_t)(v[i] & 0xFFFF);
            }
            t->num_tests++;
        }
    }
    return 0;
}

int or_task_load(const char *path, or_task_t *t)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        memset(t, 0, sizeof(*t));
        fprintf(stderr, "Cannot open task file: %s\n", path);
        return -1;
    }
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    char *json = malloc((size_t)sz + 1);
    if (!json) { fclose(f); return -1; }
    size_t n = fread(json, 1, (size_t)sz, f);
    json[n] = 0;
    fclose(f);
    int rc = or_task_parse(json, t);
    free(json);
    return rc;
}

/* tasks/spec.c:188-215 */
int or_task_evaluate(const or_task_t *t, const uint16_t *prog, int len)
{
    int passed = 0;
    for (int k = 0; k < t->num_tests; k++) {
        or_vm_t vm;
        or_vm_init(&vm);
        or_vm_load(&vm, prog, (size_t)len);
        for (int i = 0; i < t->num_inputs; i++) vm.ports[t->input_ports[i]] = t->tests[k].in[i];
        or_vm_run(&vm, t->max_cycles);
        if (vm.status != OR_HALTED) continue;
        int ok = 1;
        for (int i = 0; i < t->num_outputs; i++)
            if (vm.ports[t->output_ports[i]] != t->tests[k].out[i]) { ok = 0; break; }
        if (ok) passed++;
    }
    return passed;
}

/* pass/fail only, stopping at the first failing test: same verdict as
 * or_task_evaluate(...) == num_tests, used to keep brute force tractable */
static int or_passes(const or_task_t *t, const uint16_t *prog, int len)
{
    for (int k = 0; k < t->num_tests; k++) {
        or_vm_t vm;
        or_vm_init(&vm);
        or_vm_load(&vm, prog, (size_t)len);
        for (int i = 0; i < t->num_inputs; i++) vm.ports[t->input_ports[i]] = t->tests[k].in[i];
        or_vm_run(&vm, t->max_cycles);
        if (vm.status != OR_HALTED) return 0;
        for (int i = 0; i < t->num_outputs; i++)
            if (vm.ports[t->output_ports[i]] != t->tests[k].out[i]) return 0;
    }
    return 1;
}

/* --------------------------------------------------------- disassembler */
/* src/asm.c:404-503 */
static const char *opn[] = {"NOP", "HLT", "ADD", "SUB", "MUL", "DIV", "MOD",
                            "AND", "OR",  "XOR", "SHL", "SHR", "ASR", "CMP",
                            "LD",  "ST",  "LDI", "LUI", "ADDI", "BR", "JMP",
                            "CALL", "RET", "IN", "OUT"};
static const char *cndn[] = {"BRA", "BZ", "BNZ", "BLT", "BGE", "BLTU", "BGEU"};

void or_disasm(uint16_t insn, char *buf, size_t len)
{
    uint8_t op = I_OP(insn), rd = I_RD(insn), rs = I_RS(insn), rt = I_RT(insn);
    uint8_t imm = I_IMM(insn);
    int8_t simm = (int8_t)imm;
    if (op >= O_COUNT) { snprintf(buf, len, "??? 0x%04X", insn); return; }
    switch (op) {
    case O_NOP: snprintf(buf, len, "NOP"); break;
    case O_HLT: snprintf(buf, len, "HLT"); break;
    case O_RET: snprintf(buf, len, "RET"); break;
    case O_ADD: case O_SUB: case O_MUL: case O_DIV: case O_MOD: case O_AND:
    case O_OR: case O_XOR: case O_SHL: case O_SHR: case O_ASR:
        if (op == O_ADD && rt == 0) snprintf(buf, len, "MOV  r%d, r%d", rd, rs);
        else if (op == O_SUB && rs == 0) snprintf(buf, len, "NEG  r%d, r%d", rd, rt);
        else snprintf(buf, len, "%-4s r%d, r%d, r%d", opn[op], rd, rs, rt);
        break;
    case O_CMP: snprintf(buf, len, "CMP  r%d, r%d", rd, rs); break;
    case O_LD: snprintf(buf, len, "LD   r%d, r%d", rd, rs); break;
    case O_ST: snprintf(buf, len, "ST   r%d, r%d", rd, rs); break;
    case O_JMP: snprintf(buf, len, "JMP  r%d", rd); break;
    case O_CALL: snprintf(buf, len, "CALL r%d", rd); break;
    case O_LDI: snprintf(buf, len, "LDI  r%d, %d", rd, imm); break;
    case O_LUI: snprintf(buf, len, "LUI  r%d, 0x%02X", rd, imm); break;
    case O_ADDI: snprintf(buf, len, "ADDI r%d, %d", rd, simm); break;
    case O_BR:
        if (rd < 7) snprintf(buf, len, "%-4s %+d", cndn[rd], simm);
        else snprintf(buf, len, "BR   %d, %+d", rd, simm);
        break;
    case O_IN: snprintf(buf, len, "IN   r%d, %d", rd, imm & 7); break;
    case O_OUT: snprintf(buf, len, "OUT  r%d, %d", rd, imm & 7); break;
    default: snprintf(buf, len, "??? 0x%04X", insn); break;
    }
}

/* ------------------------------------------------------------ grammars */
static const uint8_t LDI_V[] = {0, 1, 2, 3, 4, 5, 8, 16, 32, 64, 128, 255};
static const int8_t ADDI_V[] = {-1, 1, -2, 2, -3, 3};

/* tools/enumerate.c:206-267 */
int or_gen_candidates(uint16_t *out, int max, uint8_t live, int pos,
                      int kernel_depth, int use_branches)
{
    static const uint8_t ops[] = {O_ADD, O_SUB, O_MUL, O_DIV, O_MOD, O_AND,
                                  O_OR,  O_XOR, O_SHL, O_SHR, O_ASR};
    static const int comm[] = {1, 0, 1, 0, 0, 1, 1, 1, 0, 0, 0};
    int n = 0;
    for (int oi = 0; oi < 11; oi++) {
        int op = ops[oi];
        for (int rd = 1; rd < 8 && n < max; rd++)
            for (int rs = 0; rs < 8 && n < max; rs++) {
                if (rs && !(live & (1 << rs))) continue;
                for (int rt = comm[oi] ? rs : 0; rt < 8 && n < max; rt++) {
                    if (rt && !(live & (1 << rt))) continue;
                    if (rs == 0 && rt == 0) continue;
                    if (op == O_ADD && rs == rd && rt == 0) continue;
                    if (op == O_ADD && rt == rd && rs == 0) continue;
                    if (op == O_SUB && rs == rd && rt == 0) continue;
                    if (op == O_MUL && (rs == 0 || rt == 0)) continue;
                    if ((op == O_DIV || op == O_MOD) && rt == 0) continue;
                    if ((op == O_SHL || op == O_SHR || op == O_ASR) && rt == 0) continue;
                    out[n++] = ENC_R(op, rd, rs, rt);
                }
            }
    }
    for (int rd = 0; rd < 8 && n < max; rd++) {
        if (rd && !(live & (1 << rd))) continue;
        for (int rs = 0; rs < 8 && n < max; rs++) {
            if (rs && !(live & (1 << rs))) continue;
            if (rd == 0 && rs == 0) continue;
            if (rd == rs) continue;
            out[n++] = ENC_R(O_CMP, rd, rs, 0);
        }
    }
    if (use_branches) {
        int mf = kernel_depth - pos - 1;
        if (mf > 0)
            for (int c = 1; c <= 6; c++)
                for (int off = 1; off <= mf && n < max; off++)
                    out[n++] = ENC_I(O_BR, c, (uint8_t)(int8_t)off);
    }
    for (int rd = 1; rd < 8 && n < max; rd++)
        for (int i = 0; i < 12 && n < max; i++) out[n++] = ENC_I(O_LDI, rd, LDI_V[i]);
    for (int rd = 1; rd < 8 && n < max; rd++) {
        if (!(live & (1 << rd))) continue;
        for (int i = 0; i < 6 && n < max; i++) out[n++] = ENC_I(O_ADDI, rd, (uint8_t)ADDI_V[i]);
    }
    return n;
}

/* src/search_state.c:280-330 */
int or_gen_candidates_branchless(uint16_t *out, int max, uint8_t live)
{
    static const uint8_t ops[] = {O_ADD, O_SUB, O_MUL, O_AND, O_OR,
                                  O_XOR, O_SHL, O_SHR, O_ASR};
    static const int comm[] = {1, 0, 1, 1, 1, 1, 0, 0, 0};
    int n = 0;
    for (int oi = 0; oi < 9; oi++) {
        int op = ops[oi];
        for (int rd = 1; rd < 8 && n < max; rd++)
            for (int rs = 0; rs < 8 && n < max; rs++) {
                if (rs && !(live & (1 << rs))) continue;
                for (int rt = comm[oi] ? rs : 0; rt < 8 && n < max; rt++) {
                    if (rt && !(live & (1 << rt))) continue;
                    if (rs == 0 && rt == 0) continue;
                    if (op == O_ADD && rs == rd && rt == 0) continue;
                    if (op == O_ADD && rt == rd && rs == 0) continue;
                    if (op == O_SUB && rs == rd && rt == 0) continue;
                    if (op == O_MUL && (rs == 0 || rt == 0)) continue;
                    if ((op == O_SHL || op == O_SHR || op == O_ASR) && rt == 0) continue;
                    out[n++] = ENC_R(op, rd, rs, rt);
                }
            }
    }
    for (int rd = 1; rd < 8 && n < max; rd++)
        for (int i = 0; i < 12 && n < max; i++) out[n++] = ENC_I(O_LDI, rd, LDI_V[i]);
    for (int rd = 1; rd < 8 && n < max; rd++) {
        if (!(live & (1 << rd))) continue;
        for (int i = 0; i < 6 && n < max; i++) out[n++] = ENC_I(O_ADDI, rd, (uint8_t)ADDI_V[i]);
    }
    return n;
}

/* tools/enumerate.c:427-433 */
uint8_t or_live_after(uint8_t live, uint16_t insn)
{
    uint8_t op = I_OP(insn), rd = I_RD(insn);
    if (op != O_CMP && op != O_BR && op != O_NOP && op != O_HLT && op != O_ST && rd != 0)
        live |= (uint8_t)(1u << rd);
    return live;
}

/* --------------------------------------------------------------- hashes */
uint32_t or_fnv32_bytes(uint32_t h, const uint8_t *p, size_t n)
{
    for (size_t i = 0; i < n; i++) { h ^= p[i]; h *= 16777619u; }
    return h;
}
uint64_t or_fnv64_bytes(uint64_t h, const uint8_t *p, size_t n)
{
    for (size_t i = 0; i < n; i++) { h ^= p[i]; h *= 0x00000100000001B3ULL; }
    return h;
}

/* ----------------------------------------------------- exact state sets */
typedef struct {
    uint64_t *hs;      /* 0 = empty (hashes are forced odd)                 */
    uint32_t *idx;
    size_t cap, n;
    uint8_t *arena;
    size_t arena_cap;
    int klen;
} xset_t;

static void xset_init(xset_t *s, int klen)
{
    memset(s, 0, sizeof(*s));
    s->klen = klen;
}
static void xset_free(xset_t *s)
{
    free(s->hs); free(s->idx); free(s->arena);
    memset(s, 0, sizeof(*s));
}
static void xset_clear(xset_t *s)
{
    if (s->hs) memset(s->hs, 0, s->cap * sizeof(uint64_t));
    s->n = 0;
}
static int xset_grow(xset_t *s)
{
    size_t ncap = s->cap ? s->cap * 2 : 1024;
    uint64_t *nh = calloc(ncap, sizeof(uint64_t));
    uint32_t *ni = malloc(ncap * sizeof(uint32_t));
    if (!nh || !ni) { free(nh); free(ni); return -1; }
    for (size_t i = 0; i < s->cap; i++) {
        if (!s->hs[i]) continue;
        size_t j = (size_t)(s->hs[i] >> 1) & (ncap - 1);
        while (nh[j]) j = (j + 1) & (ncap - 1);
        nh[j] = s->hs[i];
        ni[j] = s->idx[i];
    }
    free(s->hs); free(s->idx);
    s->hs = nh; s->idx = ni; s->cap = ncap;
    return 0;
}
/* returns 1 if newly inserted, 0 if present, -1 on allocation failure */
static int xset_insert(xset_t *s, const uint8_t *key, int insert)
{
    uint64_t h = or_fnv64_bytes(0xCBF29CE484222325ULL, key, (size_t)s->klen) | 1;
    if (s->cap == 0 || (s->n + 1) * 2 > s->cap)
        if (xset_grow(s)) return -1;
    size_t j = (size_t)(h >> 1) & (s->cap - 1);
    while (s->hs[j]) {
        if (s->hs[j] == h &&
            memcmp(s->arena + (size_t)s->idx[j] * s->klen, key, (size_t)s->klen) == 0)
            return 0;
        j = (j + 1) & (s->cap - 1);
    }
    if (!insert) return 1;
    if ((s->n + 1) * (size_t)s->klen > s->arena_cap) {
        size_t nc = s->arena_cap ? s->arena_cap * 2 : (size_t)s->klen * 1024;
        uint8_t *na = realloc(s->arena, nc);
        if (!na) return -1;
        s->arena = na;
        s->arena_cap = nc;
    }
    memcpy(s->arena + s->n * s->klen, key, (size_t)s->klen);
    s->hs[j] = h;
    s->idx[j] = (uint32_t)s->n;
    s->n++;
    return 1;
}

/* -------------------------------------------------------------- solver */
typedef struct {
    const or_task_t *task;
    or_opts_t o;
    or_result_t *res;
    uint16_t program[OR_PROG_SIZE];
    int num_inputs;
    int cap;                 /* solution cap (INT32_MAX if none)            */
    int nsol;
    size_t sol_alloc;
    uint64_t tried, leaf_tried, pruned, collisions, saturations;
    uint32_t *ref_tab[16];
    xset_t xs[16];
    int klen;
    /* frontier emission (or_frontier) */
    int split;
    uint16_t *fr_insns;
    uint8_t *fr_live;
    long fr_n, fr_cap;
    /* time limit */
    struct timespec t0;
    uint64_t node_ctr;
    int stop;
} solver_t;

static double now_since(const struct timespec *t0)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)(t.tv_sec - t0->tv_sec) + (double)(t.tv_nsec - t0->tv_nsec) * 1e-9;
}

static void check_time(solver_t *s)
{
    if (s->o.time_limit <= 0) return;
    if ((++s->node_ctr & 1023) == 0 && now_since(&s->t0) > s->o.time_limit) s->stop = 1;
}

/* tools/enumerate.c:54-77 plus the exact shadow; returns 1 to prune */
static int oep_check(solver_t *s, int depth, uint32_t h32, const uint8_t *key)
{
    if (s->o.oep_mode == OR_OEP_OFF) return 0;
    if (s->o.oep_mode == OR_OEP_EXACT) {
        int r = xset_insert(&s->xs[depth], key, 1);
        if (r == 0) { s->pruned++; return 1; }
        return 0;
    }
    /* REF32 */
    size_t tsize = (size_t)1 << s->o.oep_bits, tmask = tsize - 1;
    if (!s->ref_tab[depth]) {
        s->ref_tab[depth] = calloc(tsize, sizeof(uint32_t));
        if (!s->ref_tab[depth]) return 0;
    }
    uint32_t k = h32 | 1, idx = h32 & (uint32_t)tmask;
    for (int probe = 0; probe < 32; probe++) {
        uint32_t *slot = &s->ref_tab[depth][(idx + (uint32_t)probe) & tmask];
        if (*slot == 0) {
            *slot = k;
            if (s->o.shadow_exact) xset_insert(&s->xs[depth], key, 1);
            return 0;
        }
        if (*slot == k) {
            s->pruned++;
            if (s->o.shadow_exact && xset_insert(&s->xs[depth], key, 0) == 1) s->collisions++;
            return 1;
        }
    }
    s->saturations++;
    if (s->o.shadow_exact) xset_insert(&s->xs[depth], key, 1);
    return 0;
}

static void add_solution(solver_t *s, int plen)
{
    if (s->nsol >= s->cap) s->res->cap_overflow = 1;
    if ((size_t)s->nsol >= s->sol_alloc) {
        size_t na = s->sol_alloc ? s->sol_alloc * 2 : 128;
        s->res->solutions = realloc(s->res->solutions, na * sizeof(*s->res->solutions));
        s->res->solution_lens = realloc(s->res->solution_lens, na * sizeof(int));
        s->sol_alloc = na;
    }
    memset(s->res->solutions[s->nsol], 0, sizeof(s->res->solutions[0]));
    memcpy(s->res->solutions[s->nsol], s->program, (size_t)plen * sizeof(uint16_t));
    s->res->solution_lens[s->nsol] = plen;
    s->nsol++;
}

/* tools/enumerate.c:271-290 */
static void try_output(solver_t *s, uint8_t live, int kend)
{
    if (s->nsol >= s->cap) return;
    for (int r = 1; r < 8; r++) {
        if (!(live & (1 << r))) continue;
        s->program[kend] = ENC_I(O_OUT, r, s->task->output_ports[0]);
        s->program[kend + 1] = ENC_R(O_HLT, 0, 0, 0);
        s->tried++;
        s->leaf_tried++;
        if (or_task_evaluate(s->task, s->program, kend + 2) == s->task->num_tests)
            add_solution(s, kend + 2);
    }
}

/* VM replay of the partial program for OEP: tools/enumerate.c:435-452 */
static uint32_t replay_key(solver_t *s, int partial_len, uint8_t *key)
{
    const or_task_t *t = s->task;
    s->program[partial_len] = ENC_R(O_HLT, 0, 0, 0);
    uint32_t h = 2166136261u;
    uint8_t *kp = key;
    for (int k = 0; k < t->num_tests; k++) {
        or_vm_t vm;
        or_vm_init(&vm);
        or_vm_load(&vm, s->program, (size_t)partial_len + 1);
        for (int p = 0; p < t->num_inputs; p++) vm.ports[t->input_ports[p]] = t->tests[k].in[p];
        or_vm_run(&vm, 100);
        for (int r = 0; r < OR_NUM_REGS; r++) {
            *kp++ = (uint8_t)(vm.r[r] & 0xFF);
            *kp++ = (uint8_t)(vm.r[r] >> 8);
        }
        *kp++ = vm.flags;
    }
    h = or_fnv32_bytes(h, key, (size_t)(kp - key));
    return h;
}

/* tools/enumerate.c:405-461 (branch-mode DFS); split >= 0 emits prefixes
 * at pos == split instead of descending (gpu/hybrid.cu:524-588) */
static void dfs_branch(solver_t *s, int pos, int D, uint8_t live)
{
    if (s->nsol >= s->cap || s->stop) return;
    int kstart = s->num_inputs;
    if (s->split >= 0 && pos == s->split) {
        if (s->fr_n < s->fr_cap) {
            for (int i = 0; i < 8; i++)
                s->fr_insns[s->fr_n * 8 + i] = i < pos ? s->program[kstart + i] : 0;
            s->fr_live[s->fr_n] = live;
        }
        s->fr_n++;
        return;
    }
    if (pos == D) { try_output(s, live, kstart + D); return; }
    uint16_t cands[4096];
    int nc = or_gen_candidates(cands, 4096, live, pos, D, s->o.use_branches);
    int pp = kstart + pos;
    uint8_t key[OR_MAX_TESTS * 17];
    for (int i = 0; i < nc; i++) {
        if (s->nsol >= s->cap || s->stop) return;
        check_time(s);
        s->program[pp] = cands[i];
        uint8_t nl = or_live_after(live, cands[i]);
        if (s->o.oep_mode != OR_OEP_OFF && I_OP(cands[i]) != O_BR) {
            uint32_t h = replay_key(s, pp + 1, key);
            if (oep_check(s, pos, h, key)) continue;
        }
        s->tried++;
        dfs_branch(s, pos + 1, D, nl);
    }
}

/* ---- branchless mode: src/search_state.c:15-239, tools/enumerate.c:292-402 */
typedef struct {
    uint16_t regs[OR_MAX_TESTS][8];
    uint8_t flags[OR_MAX_TESTS];
    uint8_t live, depth;
} ss_t;

static int ss_step(const or_task_t *t, const ss_t *src, uint16_t insn, ss_t *dst)
{
    uint8_t op = I_OP(insn), rd = I_RD(insn), rs = I_RS(insn), rt = I_RT(insn);
    uint8_t imm = I_IMM(insn);
    switch (op) {
    case O_ADD: case O_SUB: case O_MUL: case O_DIV: case O_MOD: case O_AND:
    case O_OR: case O_XOR: case O_SHL: case O_SHR: case O_ASR: case O_LDI: case O_ADDI:
        break;
    default: return 0;
    }
    *dst = *src;
    dst->depth = (uint8_t)(src->depth + 1);
    for (int k = 0; k < t->num_tests; k++) {
        uint16_t *r = dst->regs[k];
        uint8_t *f = &dst->flags[k];
        if (op == O_LDI) {
            r[rd] = imm;
        } else if (op == O_ADDI) {
            uint16_t ext = (uint16_t)(int16_t)(int8_t)imm;
            *f = fl_add(r[rd], ext);
            r[rd] = (uint16_t)(r[rd] + ext);
        } else {
            int fault;
            uint8_t nf;
            uint16_t v = or_alu(op, r[rs], r[rt], *f, &nf, &fault);
            if (fault) return 0;
            r[rd] = v;
            *f = nf;
        }
        r[0] = 0;
    }
    if (rd != 0) dst->live = (uint8_t)(src->live | (1u << rd));
    return 1;
}

static void ss_key(const or_task_t *t, const ss_t *st, uint8_t *key)
{
    uint8_t *kp = key;
    for (int k = 0; k < t->num_tests; k++) {
        for (int r = 0; r < 8; r++) {
            *kp++ = (uint8_t)(st->regs[k][r] & 0xFF);
            *kp++ = (uint8_t)(st->regs[k][r] >> 8);
        }
        *kp++ = st->flags[k];
    }
}

static void dfs_ss(solver_t *s, const ss_t *st, int pos, int D)
{
    if (s->nsol >= s->cap || s->stop) return;
    int kstart = s->num_inputs;
    const or_task_t *t = s->task;
    if (pos == D) {
        /* tools/enumerate.c:292-313 with src/search_state.c:223-239 */
        if (s->nsol >= s->cap) return;
        uint8_t solving = 0;
        for (int r = 1; r < 8; r++) {
            if (!(st->live & (1u << r))) continue;
            int ok = 1;
            for (int k = 0; k < t->num_tests; k++)
                if (st->regs[k][r] != t->tests[k].out[0]) { ok = 0; break; }
            if (ok) solving |= (uint8_t)(1u << r);
        }
        int kend = kstart + D;
        for (int r = 1; r < 8; r++) {
            if (!(solving & (1u << r))) continue;
            s->program[kend] = ENC_I(O_OUT, r, t->output_ports[0]);
            s->program[kend + 1] = ENC_R(O_HLT, 0, 0, 0);
            s->tried++;
            s->leaf_tried++;
            if (or_task_evaluate(t, s->program, kend + 2) == t->num_tests) add_solution(s, kend + 2);
        }
        return;
    }
    uint16_t cands[4096];
    int nc = or_gen_candidates_branchless(cands, 4096, st->live);
    int pp = kstart + pos;
    uint8_t key[OR_MAX_TESTS * 17];
    for (int i = 0; i < nc; i++) {
        if (s->nsol >= s->cap || s->stop) return;
        check_time(s);
        ss_t child;
        if (!ss_step(t, st, cands[i], &child)) continue;
        s->program[pp] = cands[i];
        s->tried++;
        if (s->o.oep_mode != OR_OEP_OFF) {
            ss_key(t, &child, key);
            uint64_t h64 = or_fnv64_bytes(0xCBF29CE484222325ULL, key, (size_t)t->num_tests * 17);
            if (oep_check(s, pos, (uint32_t)h64, key)) continue;
        }
        dfs_ss(s, &child, pos + 1, D);
    }
}

void or_opts_default(or_opts_t *o)
{
    memset(o, 0, sizeof(*o));
    o->use_branches = 1;
    o->oep_mode = OR_OEP_EXACT;
    o->oep_bits = 20;
    o->max_solutions = 100;
    o->start_depth = 0;
    o->max_depth = 4;
}

void or_result_free(or_result_t *r)
{
    free(r->solutions);
    free(r->solution_lens);
    r->solutions = NULL;
    r->solution_lens = NULL;
}

static void task_print(const or_task_t *t)
{
    /* tasks/spec.c:219-237 */
    printf("Task: %s\n", t->name);
    printf("  max_cycles: %llu\n", (unsigned long long)t->max_cycles);
    printf("  inputs: ");
    for (int i = 0; i < t->num_inputs; i++) printf("port%d ", t->input_ports[i]);
    printf("\n  outputs: ");
    for (int i = 0; i < t->num_outputs; i++) printf("port%d ", t->output_ports[i]);
    printf("\n  tests: %d\n", t->num_tests);
    for (int k = 0; k < t->num_tests; k++) {
        printf("    [%d] in=(", k);
        for (int i = 0; i < t->num_inputs; i++) printf("%s%u", i ? "," : "", t->tests[k].in[i]);
        printf(") out=(");
        for (int i = 0; i < t->num_outputs; i++) printf("%s%u", i ? "," : "", t->tests[k].out[i]);
        printf(")\n");
    }
}

static void print_solution(solver_t *s, int idx)
{
    /* tools/enumerate.c:465-489 */
    const or_task_t *t = s->task;
    int len = s->res->solution_lens[idx];
    printf("  Solution #%d (%d instructions):\n", idx + 1, len);
    char buf[64];
    for (int i = 0; i < len; i++) {
        or_disasm(s->res->solutions[idx][i], buf, sizeof(buf));
        printf("    %02X: %04X  %s\n", i, s->res->solutions[idx][i], buf);
    }
    or_vm_t vm;
    or_vm_init(&vm);
    or_vm_load(&vm, s->res->solutions[idx], (size_t)len);
    for (int i = 0; i < t->num_inputs; i++) vm.ports[t->input_ports[i]] = t->tests[0].in[i];
    or_vm_run(&vm, t->max_cycles);
    printf("    verify test[0]: in=(");
    for (int i = 0; i < t->num_inputs; i++) printf("%s%u", i ? "," : "", t->tests[0].in[i]);
    printf(") -> port%d=%u (expected %u) %s\n", t->output_ports[0],
           vm.ports[t->output_ports[0]], t->tests[0].out[0],
           vm.ports[t->output_ports[0]] == t->tests[0].out[0] ? "OK" : "FAIL");
}

static void solver_setup(solver_t *s, const or_task_t *task, const or_opts_t *o, or_result_t *res)
{
    memset(s, 0, sizeof(*s));
    s->task = task;
    s->o = *o;
    s->res = res;
    s->cap = o->max_solutions > 0 ? o->max_solutions : 0x7FFFFFFF;
    s->split = -1;
    s->klen = task->num_tests * 17;
    for (int d = 0; d < 16; d++) xset_init(&s->xs[d], s->klen > 0 ? s->klen : 1);
    s->num_inputs = task->num_inputs;
    for (int i = 0; i < s->num_inputs; i++)
        s->program[i] = ENC_I(O_IN, i + 1, task->input_ports[i]);
}

static void solver_teardown(solver_t *s)
{
    for (int d = 0; d < 16; d++) { free(s->ref_tab[d]); xset_free(&s->xs[d]); }
}

static void clear_tables(solver_t *s)
{
    size_t tsize = (size_t)1 << s->o.oep_bits;
    for (int d = 0; d < 16; d++) {
        if (s->ref_tab[d]) memset(s->ref_tab[d], 0, tsize * sizeof(uint32_t));
        xset_clear(&s->xs[d]);
    }
}

/* tools/enumerate.c:537-662 (pass 1; recording is out of scope) */
int or_solve(const or_task_t *task, const or_opts_t *opts, or_result_t *res)
{
    memset(res, 0, sizeof(*res));
    res->optimal_depth = -1;
    solver_t *s = calloc(1, sizeof(solver_t));
    if (!s) return -1;
    solver_setup(s, task, opts, res);

    /* tools/enumerate.c:582-584, :610: branchless needs a single-output task */
    int use_ss = task->num_tests <= OR_MAX_TESTS && task->num_outputs == 1 &&
                 task->num_inputs <= 7;
    int branchless = use_ss && !opts->use_branches;
    uint8_t initial_live = 0;
    for (int i = 0; i < task->num_inputs; i++) initial_live |= (uint8_t)(1 << (i + 1));
    ss_t root;
    memset(&root, 0, sizeof(root));
    for (int k = 0; k < task->num_tests; k++)
        for (int i = 0; i < task->num_inputs && i < 7; i++) root.regs[k][i + 1] = task->tests[k].in[i];
    root.live = initial_live;

    if (opts->print) {
        task_print(task);
        printf("\n");
        printf("Enumerating programs (max kernel depth %d, branches %s, OEP %s%s)...\n\n",
               opts->max_depth, opts->use_branches ? "on" : "off",
               opts->oep_mode != OR_OEP_OFF ? "on" : "off", branchless ? ", search_state" : "");
    }
    clock_gettime(CLOCK_MONOTONIC, &s->t0);
    for (int depth = opts->start_depth; depth <= opts->max_depth && depth < 17; depth++) {
        int total_len = s->num_inputs + depth + 2;
        if (opts->print)
            printf("--- kernel depth %d (total program length %d) ---\n", depth, total_len);
        uint64_t tb = s->tried, lb = s->leaf_tried, pb = s->pruned, cb = s->collisions,
                 sb = s->saturations;
        int sols_before = s->nsol;
        clear_tables(s);
        if (branchless) dfs_ss(s, &root, 0, depth);
        else dfs_branch(s, 0, depth, initial_live);
        double el = now_since(&s->t0);
        res->tried[depth] = s->tried - tb;
        res->leaf_tried[depth] = s->leaf_tried - lb;
        res->pruned[depth] = s->pruned - pb;
        res->ref_collisions[depth] = s->collisions - cb;
        res->ref_saturations[depth] = s->saturations - sb;
        res->seconds = el;
        if (s->stop) { res->timed_out = 1; break; }
        res->depths_done = depth;
        if (opts->print) {
            printf("  candidates: %llu  (%.1f M/sec)\n", (unsigned long long)res->tried[depth],
                   (double)res->tried[depth] / el / 1e6);
            if (opts->oep_mode != OR_OEP_OFF)
                printf("  OEP pruned: %llu\n", (unsigned long long)s->pruned);
            printf("  solutions found: %d (total: %d)\n", s->nsol - sols_before, s->nsol);
            printf("  elapsed: %.2f sec\n\n", el);
            res->num_solutions = s->nsol;
            for (int i = sols_before; i < s->nsol; i++) print_solution(s, i);
        }
        if (s->nsol > 0) {
            res->optimal_depth = depth;
            if (opts->print) {
                printf("\nOptimal solution length: %d instructions (kernel depth %d)\n", total_len, depth);
                printf("Found %d distinct program(s) at this length.\n", s->nsol);
            }
            break;
        }
    }
    res->num_solutions = s->nsol;
    res->seconds = now_since(&s->t0);
    if (opts->print) {
        if (s->nsol == 0) {
            printf("\nNo solutions found up to kernel depth %d.\n", opts->max_depth);
        } else {
            printf("\nTotal: %llu candidates in %.2f sec", (unsigned long long)s->tried, res->seconds);
            if (opts->oep_mode != OR_OEP_OFF) printf(" (OEP pruned: %llu)", (unsigned long long)s->pruned);
            printf("\n");
        }
        fflush(stdout);
    }
    solver_teardown(s);
    free(s);
    return 0;
}

/* ------------------------------------------------------------- slices */
long or_frontier(const or_task_t *task, int D, int s_split, uint16_t *insns_out,
                 uint8_t *live_out, long cap)
{
    or_opts_t o;
    or_opts_default(&o);
    o.max_solutions = 0;
    or_result_t res;
    memset(&res, 0, sizeof(res));
    solver_t *s = calloc(1, sizeof(solver_t));
    if (!s) return -1;
    solver_setup(s, task, &o, &res);
    s->split = s_split;
    s->fr_insns = insns_out;
    s->fr_live = live_out;
    s->fr_cap = cap;
    uint8_t live = 0;
    for (int i = 0; i < task->num_inputs; i++) live |= (uint8_t)(1 << (i + 1));
    dfs_branch(s, 0, D, live);
    long n = s->fr_n;
    solver_teardown(s);
    free(s);
    or_result_free(&res);
    return n > cap ? -1 : n;
}

typedef struct {
    const or_task_t *task;
    uint16_t prog[32];
    int kstart, D, s;
    uint16_t idx[16];
    uint16_t *rows;
    long cap, n;
    uint64_t cands;
} suf_t;

static void suffix_rec(suf_t *u, int pos, uint8_t live)
{
    int pp = u->kstart + pos;
    if (pos == u->D) {
        for (int r = 1; r < 8; r++) {
            if (!(live & (1 << r))) continue;
            u->prog[pp] = ENC_I(O_OUT, r, u->task->output_ports[0]);
            u->prog[pp + 1] = ENC_R(O_HLT, 0, 0, 0);
            u->cands++;
            if (or_passes(u->task, u->prog, pp + 2)) {
                if (u->n < u->cap) {
                    int k = u->D - u->s;
                    for (int j = 0; j < k; j++) u->rows[u->n * (k + 1) + j] = u->idx[j];
                    u->rows[u->n * (k + 1) + k] = (uint16_t)r;
                }
                u->n++;
            }
        }
        return;
    }
    uint16_t cands[4096];
    int nc = or_gen_candidates(cands, 4096, live, pos, u->D, 1);
    for (int i = 0; i < nc; i++) {
        u->prog[pp] = cands[i];
        u->idx[pos - u->s] = (uint16_t)i;
        suffix_rec(u, pos + 1, or_live_after(live, cands[i]));
    }
}

long or_suffix_hits(const or_task_t *task, int D, const uint16_t *prefix, int s,
                    uint8_t live, uint16_t *rows, long cap, uint64_t *candidates_out)
{
    suf_t u;
    memset(&u, 0, sizeof(u));
    u.task = task;
    u.kstart = task->num_inputs;
    u.D = D;
    u.s = s;
    u.rows = rows;
    u.cap = cap;
    for (int i = 0; i < task->num_inputs; i++) u.prog[i] = ENC_I(O_IN, i + 1, task->input_ports[i]);
    for (int i = 0; i < s; i++) u.prog[u.kstart + i] = prefix[i];
    suffix_rec(&u, s, live);
    if (candidates_out) *candidates_out = u.cands;
    return u.n;
}

void or_replay_state(const or_task_t *task, const uint16_t *kernel, int len, uint16_t *out)
{
    uint16_t prog[40];
    int n = 0;
    for (int i = 0; i < task->num_inputs; i++) prog[n++] = ENC_I(O_IN, i + 1, task->input_ports[i]);
    for (int i = 0; i < len; i++) prog[n++] = kernel[i];
    prog[n] = ENC_R(O_HLT, 0, 0, 0);
    for (int k = 0; k < task->num_tests; k++) {
        or_vm_t vm;
        or_vm_init(&vm);
        or_vm_load(&vm, prog, (size_t)n + 1);
        for (int p = 0; p < task->num_inputs; p++) vm.ports[task->input_ports[p]] = task->tests[k].in[p];
        or_vm_run(&vm, 100);
        for (int r = 0; r < 8; r++) out[k * 9 + r] = vm.r[r];
        out[k * 9 + 8] = vm.flags;
    }
}

/* ---- the last kernel position of a depth, brute force on the VM ----
 * tools/enumerate.c:405-461 at pos == D-1 below every node of the frontier
 * at D-1: each child's replay state (scaffold + kernel + HLT, 100 cycles,
 * tools/enumerate.c:435-452), OEP on that state (first DFS visitor kept),
 * then try_output (:271-290).  Only children with a register r1..r7 equal
 * to the target on every test can pass, and two children with the same
 * state agree on that, so only those children's states are kept and the
 * first-visitor test runs among them alone.
 *
 * Each test's VM is run once per prefix up to the last slot and snapshotted
 * there; a child is the snapshot with its word in the slot, run on by the
 * same or_vm_run.  The part before the slot does not read the slot (jumps
 * are forward and land at most on the HLT), so the result is the replay's.
 */
typedef struct {
    uint32_t prefix, j, mask, flags;
} lrow_t;

typedef struct {
    lrow_t *rows;
    uint8_t *keys;
    size_t n, cap;
    uint64_t children;
} lchunk_t;

typedef struct {
    const or_task_t *task;
    int D;
    const uint16_t *insns;
    const uint8_t *live;
    long n;
    long chunk;
    long nchunks;
    lchunk_t *out;
    long next;           /* chunk counter (atomic)                         */
    int err;
} ljob_t;

static int lchunk_push(lchunk_t *c, const lrow_t *r, const uint8_t *key, int klen)
{
    if (c->n == c->cap) {
        size_t nc = c->cap ? c->cap * 2 : 256;
        lrow_t *nr = realloc(c->rows, nc * sizeof(lrow_t));
        if (!nr) return -1;
        c->rows = nr;
        uint8_t *nk = realloc(c->keys, nc * (size_t)klen);
        if (!nk) return -1;
        c->keys = nk;
        c->cap = nc;
    }
    c->rows[c->n] = *r;
    memcpy(c->keys + c->n * (size_t)klen, key, (size_t)klen);
    c->n++;
    return 0;
}

static void last_level_chunk(ljob_t *jb, long ci)
{
    const or_task_t *t = jb->task;
    const int T = t->num_tests, nin = t->num_inputs, D = jb->D, klen = T * 17;
    const int slot = nin + D - 1;
    lchunk_t *c = &jb->out[ci];
    or_vm_t *snap = malloc((size_t)T * sizeof(or_vm_t));
    if (!snap) { jb->err = 1; return; }
    or_vm_t w;
    uint8_t key[OR_MAX_TESTS * 17];
    uint16_t cands[4096];
    long i_end = (ci + 1) * jb->chunk < jb->n ? (ci + 1) * jb->chunk : jb->n;
    for (long i = ci * jb->chunk; i < i_end; i++) {
        uint16_t prog[OR_PROG_SIZE];
        int n = 0;
        for (int q = 0; q < nin; q++) prog[n++] = ENC_I(O_IN, q + 1, t->input_ports[q]);
        for (int q = 0; q < D - 1; q++) prog[n++] = jb->insns[i * 8 + q];
        prog[n++] = 0;                        /* the slot */
        prog[n++] = ENC_R(O_HLT, 0, 0, 0);
        for (int k = 0; k < T; k++) {
            or_vm_t *v = &snap[k];
            or_vm_init(v);
            or_vm_load(v, prog, (size_t)n);
            for (int p = 0; p < nin; p++) v->ports[t->input_ports[p]] = t->tests[k].in[p];
            while (v->status == OR_RUNNING && v->cycles < 100 && v->pc < slot) or_vm_step(v);
        }
        const int nc = or_gen_candidates(cands, 4096, jb->live[i], D - 1, D, 1);
        c->children += (uint64_t)nc;
        for (int j = 0; j < nc; j++) {
            uint32_t mask = 0xFE, dead = 0;
            int k;
            for (k = 0; k < T; k++) {
                w = snap[k];
                w.prog[slot] = cands[j];
                or_vm_run(&w, 100);
                uint32_t m = 0;
                for (int r = 1; r < 8; r++)
                    if (w.r[r] == t->tests[k].out[0]) m |= 1u << r;
                mask &= m;
                if (!mask) break;
                for (int r = 0; r < OR_NUM_REGS; r++) {
                    key[k * 17 + 2 * r] = (uint8_t)(w.r[r] & 0xFF);
                    key[k * 17 + 2 * r + 1] = (uint8_t)(w.r[r] >> 8);
                }
                key[k * 17 + 16] = w.flags;
            }
            if (k < T) continue;
            /* cannot pass: the whole program (OUT, HLT, the task's cycle
             * limit) does not halt on some test, whatever is output */
            prog[slot] = cands[j];
            prog[slot + 1] = ENC_I(O_OUT, 1, t->output_ports[0]);
            prog[slot + 2] = ENC_R(O_HLT, 0, 0, 0);
            for (k = 0; k < T && !dead; k++) {
                or_vm_t *v = &w;
                or_vm_init(v);
                or_vm_load(v, prog, (size_t)slot + 3);
                for (int p = 0; p < nin; p++) v->ports[t->input_ports[p]] = t->tests[k].in[p];
                if (or_vm_run(v, t->max_cycles) != OR_HALTED) dead = 1;
            }
            prog[slot] = 0;
            prog[slot + 1] = ENC_R(O_HLT, 0, 0, 0);
            lrow_t r = {(uint32_t)i, (uint32_t)j, mask, dead};
            if (lchunk_push(c, &r, key, klen)) { jb->err = 1; break; }
        }
    }
    free(snap);
}

static void *last_level_worker(void *arg)
{
    ljob_t *jb = arg;
    for (;;) {
        long ci = __atomic_fetch_add(&jb->next, 1, __ATOMIC_RELAXED);
        if (ci >= jb->nchunks || jb->err) break;
        last_level_chunk(jb, ci);
    }
    return NULL;
}

long or_last_level(const or_task_t *task, int D, const uint16_t *insns, const uint8_t *live, long n,
                   int nthreads, uint32_t *rows, long cap, uint64_t *children_out)
{
    if (D < 1 || D > 8 || n < 0) return -1;
    ljob_t jb;
    memset(&jb, 0, sizeof(jb));
    jb.task = task;
    jb.D = D;
    jb.insns = insns;
    jb.live = live;
    jb.n = n;
    jb.chunk = 2048;
    jb.nchunks = (n + jb.chunk - 1) / jb.chunk;
    jb.out = calloc((size_t)(jb.nchunks ? jb.nchunks : 1), sizeof(lchunk_t));
    if (!jb.out) return -1;
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 64) nthreads = 64;
    pthread_t th[64];
    int started = 0;
    for (int q = 1; q < nthreads; q++)
        if (pthread_create(&th[started], NULL, last_level_worker, &jb) == 0) started++;
    last_level_worker(&jb);
    for (int q = 0; q < started; q++) pthread_join(th[q], NULL);
    long total = 0;
    uint64_t children = 0;
    const int klen = task->num_tests * 17;
    xset_t xs;
    xset_init(&xs, klen);
    for (long ci = 0; ci < jb.nchunks && !jb.err; ci++) {
        lchunk_t *c = &jb.out[ci];
        children += c->children;
        for (size_t q = 0; q < c->n; q++) {
            int ins = xset_insert(&xs, c->keys + q * (size_t)klen, 1);
            if (ins < 0) { jb.err = 1; break; }
            if (total < cap) {
                uint32_t *o = rows + total * 4;
                o[0] = c->rows[q].prefix;
                o[1] = c->rows[q].j;
                o[2] = c->rows[q].mask;
                o[3] = c->rows[q].flags | (ins == 1 ? 2u : 0u);
            }
            total++;
        }
    }
    for (long ci = 0; ci < jb.nchunks; ci++) { free(jb.out[ci].rows); free(jb.out[ci].keys); }
    free(jb.out);
    xset_free(&xs);
    if (children_out) *children_out = children;
    return jb.err ? -1 : total;
}

/* ---- exhaustive ALU digest (pins the VM's ALU semantics) ----
 * Sum over a in [0, 2^16) and b in [b_lo, b_hi) of a 64-bit mix of
 * (a, b, value, flags-after, fault) with flags-in 0.  Order independent,
 * so any other implementation can produce the same figure in any order. */
static uint64_t dg_mix(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

uint64_t or_alu_digest(int op, uint32_t b_lo, uint32_t b_hi)
{
    uint64_t h = 0;
    for (uint32_t b = b_lo; b < b_hi; b++)
        for (uint32_t a = 0; a < 65536; a++) {
            uint8_t f;
            int fault;
            uint16_t v = or_alu(op, (uint16_t)a, (uint16_t)b, 0, &f, &fault);
            if (fault) { v = 0; f = 0; }
            h += dg_mix((uint64_t)a | ((uint64_t)b << 16) | ((uint64_t)v << 32) | ((uint64_t)f << 48) |
                        ((uint64_t)fault << 56));
        }
    return h;
}

/* ------------------------------------------------- exhaustive completion */
/* src/search_state.c:223-239: the live registers r1..r7 equal to the target
 * on every test. */
static uint8_t ss_leaf(const or_task_t *t, const ss_t *st)
{
    uint8_t out = 0;
    for (int r = 1; r < 8; r++) {
        if (!(st->live & (1u << r))) continue;
        int ok = 1;
        for (int k = 0; k < t->num_tests && ok; k++)
            if (st->regs[k][r] != t->tests[k].out[0]) ok = 0;
        if (ok) out |= (uint8_t)(1u << r);
    }
    return out;
}

/* src/search_state.c:372-415: DFS to `target` appended instructions; the
 * leaf check runs on every child, the OEP table of the parent position on
 * children that still have room below them. */
static int complete_dfs(solver_t *s, const ss_t *st, int pos, int target, uint16_t *path)
{
    const or_task_t *t = s->task;
    uint16_t cands[4096];
    uint8_t key[OR_MAX_TESTS * 17];
    int nc = or_gen_candidates_branchless(cands, 4096, st->live);
    for (int i = 0; i < nc; i++) {
        ss_t child;
        if (!ss_step(t, st, cands[i], &child)) continue;
        s->tried++;
        if (ss_leaf(t, &child)) {
            path[pos] = cands[i];
            return pos + 1;
        }
        if (pos + 1 >= target) continue;
        ss_key(t, &child, key);
        uint64_t h64 = or_fnv64_bytes(0xCBF29CE484222325ULL, key, (size_t)t->num_tests * 17);
        if (oep_check(s, pos, (uint32_t)h64, key)) continue;
        path[pos] = cands[i];
        int len = complete_dfs(s, &child, pos + 1, target, path);
        if (len > 0) return len;
    }
    return 0;
}

/* src/search_state.c:418-445 */
int or_complete(const or_task_t *task, const or_state_t *start, int max_depth, int oep_mode,
                uint16_t *insns, int *len, uint64_t *explored)
{
    *len = 0;
    if (explored) *explored = 0;
    ss_t root;
    memcpy(&root, start, sizeof(root));
    if (ss_leaf(task, &root)) return 1;
    int max_r = max_depth - (int)root.depth;
    if (max_r <= 0) return 0;
    if (max_r > 16) max_r = 16;
    or_opts_t o;
    or_opts_default(&o);
    o.oep_mode = oep_mode;
    o.oep_bits = 20;
    or_result_t res;
    memset(&res, 0, sizeof(res));
    solver_t *s = calloc(1, sizeof(solver_t));
    if (!s) return -1;
    solver_setup(s, task, &o, &res);
    uint16_t path[16];
    int found = 0;
    for (int target = 1; target <= max_r && !found; target++) {
        clear_tables(s);
        int n = complete_dfs(s, &root, 0, target, path);
        if (n > 0) {
            memcpy(insns, path, (size_t)n * sizeof(uint16_t));
            *len = n;
            found = 1;
        }
    }
    if (explored) *explored = s->tried;
    solver_teardown(s);
    free(s);
    return found;
}

/* src/search_state.c:34-46 */
void or_state_root(const or_task_t *task, or_state_t *st)
{
    memset(st, 0, sizeof(*st));
    for (int k = 0; k < task->num_tests; k++)
        for (int i = 0; i < task->num_inputs && i < 7; i++) st->regs[k][i + 1] = task->tests[k].in[i];
    for (int i = 0; i < task->num_inputs && i < 7; i++) st->live |= (uint8_t)(1u << (i + 1));
}

int or_state_step(const or_task_t *task, const or_state_t *src, uint16_t insn, or_state_t *dst)
{
    ss_t a, b;
    memcpy(&a, src, sizeof(a));
    if (!ss_step(task, &a, insn, &b)) return 0;
    memcpy(dst, &b, sizeof(b));
    return 1;
}

uint8_t or_state_leaf(const or_task_t *task, const or_state_t *st)
{
    ss_t a;
    memcpy(&a, st, sizeof(a));
    return ss_leaf(task, &a);
}

typedef char or_state_layout_check[(sizeof(or_state_t) == 546 && sizeof(ss_t) == sizeof(or_state_t)) ? 1 : -1];
