// Checks step_packed (the register form the kernels use) against step_test
// on random states for every instruction the search grammar can place:
// ALU / CMP / LDI / ADDI over all register fields and BR with every
// condition and offset, at every kernel position.  Prints "ok <cases>" or
// the first mismatch and exits 1.
#include <cstdio>
#include <cstring>
#include <cstdint>
#include "../../pca_amd/csrc/isa.h"
using namespace pca;

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static uint32_t next32()
{
    rng ^= rng << 13; rng ^= rng >> 7; rng ^= rng << 17;
    return (uint32_t)(rng >> 16);
}

int main()
{
    static_assert(sizeof(TestState) == sizeof(PackedState), "layouts");
    const uint16_t special[] = {0, 1, 2, 0x7FFF, 0x8000, 0x8001, 0xFFFF, 0xFFFE, 15, 16};
    long cases = 0;
    for (int trial = 0; trial < 4000; trial++) {
        TestState s;
        for (int r = 0; r < 8; r++)
            s.r[r] = (next32() & 3) ? (uint16_t)next32() : special[next32() % 10];
        s.r[0] = (trial & 7) ? 0 : s.r[0];   // mostly r0 = 0, sometimes not
        s.flags = (uint8_t)(next32() & 15);
        s.resume = (uint8_t)((next32() & 3) ? 0 : next32() % 10);
        s.frozen = (uint8_t)((next32() % 16) == 0);
        s.exec = (uint8_t)next32();
        const uint32_t pos = next32() % 9;
        for (uint32_t op = 0; op < OP_COUNT; op++) {
            const bool alu = op >= OP_ADD && op <= OP_ASR;
            if (!alu && op != OP_CMP && op != OP_LDI && op != OP_ADDI && op != OP_BR) continue;
            for (uint32_t rd = 0; rd < 8; rd++)
                for (uint32_t x = 0; x < 64; x++) {
                    uint16_t insn;
                    if (alu || op == OP_CMP) insn = enc_r(op, rd, x >> 3, x & 7);
                    else if (op == OP_BR) insn = enc_i(op, rd, (x % 8) + 1);
                    else insn = enc_i(op, rd, (uint32_t)(next32() & 0xFF));
                    TestState a = s;
                    PackedState p;
                    std::memcpy(&p, &s, sizeof(p));
                    PackedState pn = p, pc = p;
                    step_test(a, insn, pos);
                    step_packed(p, insn, pos);
                    step_packed<false>(pn, insn, pos);
                    // the caller-checked form, as the kernels use it
                    if (ps_runs(pc.w[4], pos)) step_packed<false, false>(pc, insn, pos);
                    cases++;
                    if (std::memcmp(&a, &p, sizeof(p)) != 0) {
                        std::printf("mismatch insn %04x pos %u trial %d\n", insn, pos, trial);
                        return 1;
                    }
                    // the flag-free form: the same state with the flags byte untouched
                    TestState an = a;
                    an.flags = s.flags;
                    if (std::memcmp(&an, &pn, sizeof(pn)) != 0) {
                        std::printf("flag-free mismatch insn %04x pos %u trial %d\n", insn, pos, trial);
                        return 1;
                    }
                    if (std::memcmp(&pc, &pn, sizeof(pn)) != 0) {
                        std::printf("caller-checked mismatch insn %04x pos %u trial %d\n", insn, pos, trial);
                        return 1;
                    }
                }
        }
    }
    std::printf("ok %ld\n", cases);
    return 0;
}
