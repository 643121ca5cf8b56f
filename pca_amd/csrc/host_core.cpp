// host_core.cpp -- task loading, full VM, disassembler, candidate grammar.
#include "host_core.h"

#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <map>
#include <string>

namespace pca {

// ------------------------------------------------------------------ VM
void vm_init(Vm &vm) { std::memset(&vm, 0, sizeof(vm)); }

// src/vm.c:58-205
void vm_step(Vm &vm)
{
    if (vm.status != VM_RUNNING) return;
    const uint32_t insn = vm.prog[vm.pc];
    const uint32_t op = insn_op(insn), rd = insn_rd(insn), rs = insn_rs(insn);
    const uint32_t imm = insn_imm(insn);
    vm.pc = (uint8_t)(vm.pc + 1);
    if (op >= OP_ADD && op <= OP_ASR) {
        uint32_t f;
        bool fault;
        uint32_t v = alu_eval(op, vm.r[rs], vm.r[insn_rt(insn)], imm, vm.flags, &f, &fault);
        if (fault) vm.status = VM_FAULT;
        else { vm.r[rd] = (uint16_t)v; vm.flags = (uint8_t)f; }
    } else {
        switch (op) {
        case OP_NOP: break;
        case OP_HLT: vm.status = VM_HALTED; break;
        case OP_CMP: vm.flags = (uint8_t)flags_sub(vm.r[rd], vm.r[rs]); break;
        case OP_LD: vm.r[rd] = vm.mem[vm.r[rs] & 0xFF]; break;
        case OP_ST: vm.mem[vm.r[rd] & 0xFF] = vm.r[rs]; break;
        case OP_LDI: vm.r[rd] = (uint16_t)imm; break;
        case OP_LUI: vm.r[rd] = (uint16_t)((vm.r[rd] & 0xFF) | (imm << 8)); break;
        case OP_ADDI: {
            uint32_t f;
            bool fault;
            vm.r[rd] = (uint16_t)alu_eval(OP_ADDI, vm.r[rd], 0, imm, vm.flags, &f, &fault);
            vm.flags = (uint8_t)f;
            break;
        }
        case OP_BR:
            if (cond_taken(rd, vm.flags)) vm.pc = (uint8_t)((int)vm.pc + insn_simm(insn));
            break;
        case OP_JMP: vm.pc = (uint8_t)vm.r[rd]; break;
        case OP_CALL:
            if (vm.sp >= 16) { vm.status = VM_FAULT; break; }
            vm.stack[vm.sp++] = vm.pc;
            vm.pc = (uint8_t)vm.r[rd];
            break;
        case OP_RET:
            if (vm.sp == 0) { vm.status = VM_FAULT; break; }
            vm.pc = (uint8_t)vm.stack[--vm.sp];
            break;
        case OP_IN: vm.r[rd] = vm.ports[imm & 7]; break;
        case OP_OUT: vm.ports[imm & 7] = vm.r[rd]; break;
        default: vm.status = VM_FAULT; break;
        }
    }
    vm.r[0] = 0;
    vm.cycles++;
}

int vm_run(Vm &vm, uint64_t max_cycles)
{
    while (vm.status == VM_RUNNING && vm.cycles < max_cycles) vm_step(vm);
    if (vm.status == VM_RUNNING) vm.status = VM_TIMEOUT;
    return vm.status;
}

// ---------------------------------------------------------- task loader
// tasks/spec.c:64-184: a key-search reader for the task schema, keeping its
// behaviour: keys are found by substring search from the given point,
// missing max_cycles means 1000, at most 32 tests and 8 ports are read,
// values are reduced mod 2^16.
namespace {
const char *skip_space(const char *p)
{
    while (*p && std::isspace((unsigned char)*p)) ++p;
    return p;
}
const char *after_key(const char *from, const char *key)
{
    std::string pat = std::string("\"") + key + "\"";
    const char *p = std::strstr(from, pat.c_str());
    if (!p) return nullptr;
    p = skip_space(p + pat.size());
    if (*p == ':') ++p;
    return skip_space(p);
}
const char *read_long(const char *p, long *v)
{
    char *end = nullptr;
    *v = std::strtol(skip_space(p), &end, 10);
    return end;
}
int read_longs(const char *p, long *vals, int maxn)
{
    p = skip_space(p);
    if (*p != '[') return 0;
    ++p;
    int n = 0;
    while (n < maxn) {
        p = skip_space(p);
        if (*p == ']') return n;
        if (*p == ',') ++p;
        p = read_long(p, &vals[n]);
        ++n;
    }
    return n;
}
}  // namespace

int task_parse(const char *json, pca_task_t *t)
{
    std::memset(t, 0, sizeof(*t));
    if (const char *p = after_key(json, "name")) {
        if (*p == '"') {
            ++p;
            size_t i = 0;
            while (*p && *p != '"' && i + 1 < sizeof(t->name)) t->name[i++] = *p++;
        }
    }
    t->max_cycles = 1000;
    if (const char *p = after_key(json, "max_cycles")) {
        long v;
        read_long(p, &v);
        t->max_cycles = (uint64_t)v;
    }
    long vals[PCAGPU_MAX_PORTS];
    if (const char *p = after_key(json, "input_ports")) {
        t->num_inputs = read_longs(p, vals, PCAGPU_MAX_PORTS);
        for (int i = 0; i < t->num_inputs; i++) t->input_ports[i] = (uint8_t)vals[i];
    }
    if (const char *p = after_key(json, "output_ports")) {
        t->num_outputs = read_longs(p, vals, PCAGPU_MAX_PORTS);
        for (int i = 0; i < t->num_outputs; i++) t->output_ports[i] = (uint8_t)vals[i];
    }
    const char *p = after_key(json, "tests");
    if (!p) return 0;
    p = skip_space(p);
    if (*p == '[') ++p;
    while (t->num_tests < PCAGPU_MAX_TESTS) {
        p = skip_space(p);
        if (*p == ']' || *p == 0) break;
        if (*p == ',') ++p;
        p = skip_space(p);
        if (*p != '{') break;
        const char *obj = p;
        int depth = 1;
        for (++p; *p && depth > 0; ++p) depth += (*p == '{') - (*p == '}');
        pca_test_case_t &tc = t->tests[t->num_tests];
        if (const char *q = after_key(obj, "in")) {
            int n = read_longs(q, vals, PCAGPU_MAX_PORTS);
            for (int i = 0; i < n; i++) tc.in[i] = (uint16_t)(vals[i] & 0xFFFF);
        }
        if (const char *q = after_key(obj, "out")) {
            int n = read_longs(q, vals, PCAGPU_MAX_PORTS);
            for (int i = 0; i < n; i++) tc.out[i] = (uint16_t)(vals[i] & 0xFFFF);
        }
        t->num_tests++;
    }
    return 0;
}

int task_load(const char *path, pca_task_t *t)
{
    std::memset(t, 0, sizeof(*t));
    FILE *f = std::fopen(path, "rb");
    if (!f) return PCAGPU_EIO;
    std::string s;
    char buf[4096];
    size_t n;
    while ((n = std::fread(buf, 1, sizeof(buf), f)) > 0) s.append(buf, n);
    std::fclose(f);
    return task_parse(s.c_str(), t);
}

// tasks/spec.c:188-215
int task_evaluate(const pca_task_t *t, const uint16_t *prog, int len)
{
    int passed = 0;
    if (len > 256) len = 256;
    for (int k = 0; k < t->num_tests; k++) {
        Vm vm;
        vm_init(vm);
        if (len > 0) std::memcpy(vm.prog, prog, (size_t)len * 2);
        for (int i = 0; i < t->num_inputs; i++) vm.ports[t->input_ports[i] & 7] = t->tests[k].in[i];
        vm_run(vm, t->max_cycles);
        if (vm.status != VM_HALTED) continue;
        bool ok = true;
        for (int i = 0; i < t->num_outputs && ok; i++) ok = vm.ports[t->output_ports[i] & 7] == t->tests[k].out[i];
        passed += ok;
    }
    return passed;
}

// --------------------------------------------------------- disassembler
// src/asm.c:404-503: mnemonics, MOV/NEG pseudo-ops and column layout.
void disasm(uint16_t insn, char *buf, size_t len)
{
    static const char *const names[OP_COUNT] = {
        "NOP", "HLT", "ADD", "SUB", "MUL", "DIV", "MOD", "AND", "OR", "XOR", "SHL", "SHR", "ASR",
        "CMP", "LD", "ST", "LDI", "LUI", "ADDI", "BR", "JMP", "CALL", "RET", "IN", "OUT"};
    static const char *const conds[7] = {"BRA", "BZ", "BNZ", "BLT", "BGE", "BLTU", "BGEU"};
    const uint32_t op = insn_op(insn), rd = insn_rd(insn), rs = insn_rs(insn), rt = insn_rt(insn);
    const uint32_t imm = insn_imm(insn);
    if (op >= OP_COUNT) { std::snprintf(buf, len, "??? 0x%04X", insn); return; }
    if (op >= OP_ADD && op <= OP_ASR) {
        if (op == OP_ADD && rt == 0) std::snprintf(buf, len, "MOV  r%u, r%u", rd, rs);
        else if (op == OP_SUB && rs == 0) std::snprintf(buf, len, "NEG  r%u, r%u", rd, rt);
        else std::snprintf(buf, len, "%-4s r%u, r%u, r%u", names[op], rd, rs, rt);
        return;
    }
    switch (op) {
    case OP_NOP: case OP_HLT: case OP_RET: std::snprintf(buf, len, "%s", names[op]); break;
    case OP_CMP: case OP_LD: case OP_ST: std::snprintf(buf, len, "%-4s r%u, r%u", names[op], rd, rs); break;
    case OP_JMP: case OP_CALL: std::snprintf(buf, len, "%-4s r%u", names[op], rd); break;
    case OP_LDI: std::snprintf(buf, len, "LDI  r%u, %u", rd, imm); break;
    case OP_LUI: std::snprintf(buf, len, "LUI  r%u, 0x%02X", rd, imm); break;
    case OP_ADDI: std::snprintf(buf, len, "ADDI r%u, %d", rd, insn_simm(insn)); break;
    case OP_BR:
        if (rd < 7) std::snprintf(buf, len, "%-4s %+d", conds[rd], insn_simm(insn));
        else std::snprintf(buf, len, "BR   %u, %+d", rd, insn_simm(insn));
        break;
    case OP_IN: case OP_OUT: std::snprintf(buf, len, "%-4s r%u, %u", names[op], rd, imm & 7); break;
    default: std::snprintf(buf, len, "??? 0x%04X", insn); break;
    }
}

// ------------------------------------------------------------- grammar
namespace {
const uint8_t (&kLdi)[12] = kLdiImms;
const int8_t kAddi[6] = {-1, 1, -2, 2, -3, 3};
const uint32_t kAlu[11] = {OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_MOD, OP_AND,
                           OP_OR,  OP_XOR, OP_SHL, OP_SHR, OP_ASR};

bool commutative(uint32_t op) { return op == OP_ADD || op == OP_MUL || op == OP_AND || op == OP_OR || op == OP_XOR; }

// Dead-instruction rules of the reference grammar (tools/enumerate.c:226-232).
bool alu_redundant(uint32_t op, uint32_t rd, uint32_t rs, uint32_t rt)
{
    if (rs == 0 && rt == 0) return true;
    switch (op) {
    case OP_ADD: return (rt == 0 && rs == rd) || (rs == 0 && rt == rd);
    case OP_SUB: return rt == 0 && rs == rd;
    case OP_MUL: return rs == 0 || rt == 0;
    case OP_DIV: case OP_MOD: case OP_SHL: case OP_SHR: case OP_ASR: return rt == 0;
    default: return false;
    }
}

// Appends the non-BR part of the list: ALU block, CMP block -> returns the
// split point; then LDI and ADDI blocks.  branchless: the search-state
// grammar (src/search_state.c:280-330), the same blocks in the same order
// without DIV, MOD and CMP.
int build_list(uint16_t *out, uint32_t live, int *split, bool branchless = false)
{
    auto readable = [live](uint32_t r) { return r == 0 || (live >> r & 1u); };
    int n = 0;
    for (uint32_t op : kAlu) {
        if (branchless && (op == OP_DIV || op == OP_MOD)) continue;
        for (uint32_t rd = 1; rd < 8; rd++)
            for (uint32_t rs = 0; rs < 8; rs++) {
                if (!readable(rs)) continue;
                for (uint32_t rt = commutative(op) ? rs : 0; rt < 8; rt++)
                    if (readable(rt) && !alu_redundant(op, rd, rs, rt)) out[n++] = enc_r(op, rd, rs, rt);
            }
    }
    for (uint32_t rd = 0; rd < 8 && !branchless; rd++) {
        if (!readable(rd)) continue;
        for (uint32_t rs = 0; rs < 8; rs++)
            if (readable(rs) && rs != rd) out[n++] = enc_r(OP_CMP, rd, rs, 0);
    }
    *split = n;
    for (uint32_t rd = 1; rd < 8; rd++)
        for (uint8_t v : kLdi) out[n++] = enc_i(OP_LDI, rd, v);
    for (uint32_t rd = 1; rd < 8; rd++)
        if (live >> rd & 1u)
            for (int8_t v : kAddi) out[n++] = enc_i(OP_ADDI, rd, (uint8_t)v);
    return n;
}
}  // namespace

int gen_candidates_branchless(uint16_t *out, int max, uint32_t live)
{
    uint16_t tmp[CAND_LMAX];
    int split;
    int n = build_list(tmp, live & 0xFEu, &split, true);
    int k = 0;
    for (int j = 0; j < n && k < max; j++) out[k++] = tmp[j];
    return k;
}

int gen_candidates(uint16_t *out, int max, uint32_t live, int pos, int kernel_depth)
{
    uint16_t tmp[CAND_LMAX];
    int split;
    int n = build_list(tmp, live & 0xFEu, &split);
    int mf = kernel_depth - pos - 1;
    int k = 0;
    for (int j = 0; j < split && k < max; j++) out[k++] = tmp[j];
    if (mf > 0)
        for (int c = 1; c <= 6; c++)
            for (int off = 1; off <= mf && k < max; off++) out[k++] = enc_i(OP_BR, c, off);
    for (int j = split; j < n && k < max; j++) out[k++] = tmp[j];
    return k;
}

void CandTable::build(bool branchless)
{
    list.assign((size_t)128 * CAND_LMAX, 0);
    vg_insn.assign((size_t)128 * VG_MAX, 0);
    vg_rdmask.assign((size_t)128 * VG_MAX, 0);
    vg_j.assign((size_t)128 * VG_MAX * 8, 0xFFFF);
    for (int m = 0; m < 128; m++) {
        const uint32_t live = (uint32_t)m << 1;
        uint16_t *L = &list[(size_t)m * CAND_LMAX];
        n[m] = build_list(L, live, &a[m], branchless);
        cand_total[m] = 0;
        std::vector<uint16_t> keys;
        for (int j = 0; j < n[m]; j++) {
            const uint32_t c = L[j];
            cand_total[m] += (uint32_t)__builtin_popcount(live_after(live, c) & 0xFEu);
            const uint32_t op = insn_op(c), rd = insn_rd(c);
            if (op == OP_CMP) continue;
            // the value written depends on the opcode and sources only,
            // except ADDI, whose source is rd itself
            const uint16_t key = op == OP_ADDI ? (uint16_t)c : (uint16_t)(c & ~(7u << 8));
            size_t g = 0;
            while (g < keys.size() && keys[g] != key) g++;
            if (g == keys.size()) {
                keys.push_back(key);
                vg_insn[(size_t)m * VG_MAX + g] = key;
            }
            vg_rdmask[(size_t)m * VG_MAX + g] |= (uint8_t)(1u << rd);
            vg_j[((size_t)m * VG_MAX + g) * 8 + rd] = (uint16_t)j;
        }
        vg_n[m] = (int)keys.size();
    }
    bits_init.assign((size_t)128 * NBK, 0);
    bits_gidx.assign((size_t)128 * NBK * 64, 0xFFFF);

    for (int m = 0; m < 128; m++)
        for (int g = 0; g < vg_n[m]; g++) {
            const uint32_t c = vg_insn[(size_t)m * VG_MAX + g];
            const uint32_t op = insn_op(c);
            int kind, bit;
            if (op >= OP_ADD && op <= OP_ASR) {
                kind = (int)(op - OP_ADD);
                bit = (int)(insn_rs(c) * 8 + insn_rt(c));
            } else if (op == OP_ADDI) {
                const int v = insn_simm(c);
                kind = 11;
                bit = (int)insn_rd(c) * 8 + (v < 0 ? 2 * (-v - 1) : 2 * (v - 1) + 1);
            } else if (op == OP_LDI) {
                kind = 12;
                bit = 0;
                while (bit < 12 && kLdi[bit] != insn_imm(c)) bit++;
            } else {
                std::abort();   // the grammar has no other value-writing kind
            }
            uint64_t &init = bits_init[(size_t)m * NBK + kind];
            if (bit >= 64 || (init >> bit & 1ull)) std::abort();   // one group per bit
            init |= 1ull << bit;
            bits_gidx[((size_t)m * NBK + kind) * 64 + bit] = (uint16_t)g;
            // the device takes r1..r7 as the rd values of every DIV / MOD group
            if ((op == OP_DIV || op == OP_MOD) && vg_rdmask[(size_t)m * VG_MAX + g] != 0xFE) std::abort();
        }
}

std::vector<int> device_test_order(const pca_task_t *t)
{
    const int T = t->num_tests;
    auto score = [&](int k) {
        int sc = 0;
        for (int i = 0; i < t->num_inputs; i++) {
            const uint32_t x = t->tests[k].in[i];
            const int pc = __builtin_popcount(x);
            sc += 64 - 8 * std::abs(pc - 8) + (int)(std::min<uint32_t>(x, 65536u - x) >> 10);
        }
        return sc;
    };
    std::map<uint16_t, std::vector<int>> by_val;
    for (int k = 0; k < T; k++) by_val[t->tests[k].out[0]].push_back(k);
    std::vector<std::vector<int>> groups;
    for (auto &kv : by_val) {
        std::vector<int> g = kv.second;
        std::stable_sort(g.begin(), g.end(), [&](int x, int y) { return score(x) > score(y); });
        groups.push_back(g);
    }
    std::stable_sort(groups.begin(), groups.end(), [&](const std::vector<int> &x, const std::vector<int> &y) {
        return score(x[0]) > score(y[0]);
    });
    std::vector<int> order;
    for (size_t r = 0; order.size() < (size_t)T; r++)
        for (auto &g : groups)
            if (r < g.size()) order.push_back(g[r]);
    return order;
}

const CandTable &shared_grammar()
{
    static const CandTable table = [] {
        CandTable t;
        t.build(false);
        return t;
    }();
    return table;
}

const CandTable &shared_grammar_branchless()
{
    static const CandTable table = [] {
        CandTable t;
        t.build(true);
        return t;
    }();
    return table;
}

void root_states(const pca_task_t *t, TestState *out)
{
    for (int k = 0; k < t->num_tests; k++) {
        Vm vm;
        vm_init(vm);
        int n = 0;
        for (int i = 0; i < t->num_inputs; i++) vm.prog[n++] = enc_i(OP_IN, (uint32_t)i + 1, t->input_ports[i]);
        vm.prog[n] = enc_r(OP_HLT, 0, 0, 0);
        for (int i = 0; i < t->num_inputs; i++) vm.ports[t->input_ports[i] & 7] = t->tests[k].in[i];
        vm_run(vm, 100);
        TestState &s = out[k];
        std::memcpy(s.r, vm.r, sizeof(s.r));
        s.flags = vm.flags;
        s.resume = 0;
        s.frozen = 0;
        s.exec = (uint8_t)t->num_inputs;
    }
}

void oep_key_from_states(const TestState *st, int T, uint8_t *key)
{
    for (int k = 0; k < T; k++) {
        for (int r = 0; r < 8; r++) {
            *key++ = (uint8_t)(st[k].r[r] & 0xFF);
            *key++ = (uint8_t)(st[k].r[r] >> 8);
        }
        *key++ = st[k].flags;
    }
}

}  // namespace pca
