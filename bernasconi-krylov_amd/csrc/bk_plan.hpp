// bk_plan.hpp -- compile-time tile plans for the LDS-tiled sigma_x sweeps.
//
// A sweep kernel moves one tile of 2^TB amplitudes per workgroup iteration.
// The tile's local index u has TB bits; the 2^(TB-RB) threads and the 2^RB
// registers each thread holds cover those bits.  A *phase* fixes which local
// bits sit in the register index (reg[0..RB)) and which in the thread index
// (thr[0..TB-RB)); butterflies on register bits need no communication.  Between
// phases the tile is exchanged through LDS so that another set of local bits
// lands in registers.  'act' marks the register bits whose qubits are rotated
// in that phase.
//
// Two tile shapes:
//  * LOW  sweep: local bit b == global bit b (a contiguous 2^TB block).  All
//    TB bits are active.  Phase 0 holds the top RB bits in registers so that
//    the 64 lanes of a wave load 64 consecutive amplitudes.
//  * HIGH sweep: local bits [0,R) are the R lowest global bits ("rows",
//    2^R contiguous amplitudes, not rotated); local bits [R,TB) map to global
//    bits [h, h+TB-R) and are all rotated.  Lanes always take the row bits
//    first, so every load and store touches whole 2^R-amplitude rows.
#pragma once

namespace bk {

constexpr int kThreads = 256;     // workgroup size of the LOW sweep and the small kernels
constexpr int kThreadBits = 8;
constexpr int kMaxPhases = 4;
constexpr int kMaxBits = 10;      // register or thread bits per phase

struct Phase {
    int reg[kMaxBits];
    int thr[kMaxBits];
    unsigned act;  // bitmask over register index bits
};

struct Plan {
    int TB, RB, R, NP;
    bool low;
    Phase ph[kMaxPhases];
    constexpr int TT() const { return TB - RB; }  // thread bits: 2^TT threads per workgroup
};

// local-bit set helpers (TB <= 16)
constexpr Phase make_phase(const int* regbits, int RB, unsigned act_mask_local, int TB)
{
    Phase p{};
    unsigned used = 0;
    for (int i = 0; i < RB; ++i) {
        p.reg[i] = regbits[i];
        used |= 1u << regbits[i];
    }
    for (int i = RB; i < kMaxBits; ++i) p.reg[i] = -1;
    int k = 0;
    for (int b = 0; b < TB; ++b)
        if (!(used & (1u << b))) p.thr[k++] = b;
    for (; k < kMaxBits; ++k) p.thr[k] = -1;
    p.act = 0;
    for (int i = 0; i < RB; ++i)
        if (act_mask_local & (1u << regbits[i])) p.act |= 1u << i;
    return p;
}

// Plan for a LOW sweep over TB contiguous bits with RB register bits.
constexpr Plan make_low_plan(int TB, int RB)
{
    Plan pl{};
    pl.TB = TB;
    pl.RB = RB;
    pl.R = 0;
    pl.low = true;
    int np = 0;
    unsigned done = 0;
    {
        int regs[kMaxBits] = {};
        for (int i = 0; i < RB; ++i) regs[i] = TB - RB + i;
        unsigned act = 0;
        for (int i = 0; i < RB; ++i) act |= 1u << regs[i];
        pl.ph[np++] = make_phase(regs, RB, act, TB);
        done |= act;
    }
    int next = 0;  // remaining bits [0, TB-RB), chunked from the bottom
    while (next < TB - RB) {
        int regs[kMaxBits] = {};
        int k = 0;
        unsigned act = 0;
        while (k < RB && next < TB - RB) {
            regs[k++] = next;
            act |= 1u << next;
            ++next;
        }
        // filler: already-rotated bits from the top
        for (int b = TB - 1; k < RB && b >= 0; --b)
            if ((done & (1u << b)) && !(act & (1u << b))) regs[k++] = b;
        // keep register bits ascending
        for (int i = 0; i < RB; ++i)
            for (int j = i + 1; j < RB; ++j)
                if (regs[j] < regs[i]) {
                    int t = regs[i];
                    regs[i] = regs[j];
                    regs[j] = t;
                }
        pl.ph[np++] = make_phase(regs, RB, act, TB);
        done |= act;
    }
    pl.NP = np;
    return pl;
}

// Plan for a HIGH sweep: R row bits, TB-R active bits.
constexpr Plan make_high_plan(int TB, int RB, int R)
{
    Plan pl{};
    pl.TB = TB;
    pl.RB = RB;
    pl.R = R;
    pl.low = false;
    int np = 0;
    unsigned done = 0;
    int next = R;
    while (next < TB) {
        int regs[kMaxBits] = {};
        int k = 0;
        unsigned act = 0;
        while (k < RB && next < TB) {
            regs[k++] = next;
            act |= 1u << next;
            ++next;
        }
        for (int b = TB - 1; k < RB && b >= R; --b)
            if ((done & (1u << b)) && !(act & (1u << b))) regs[k++] = b;
        // still short (fewer active bits than RB): borrow the highest row bits
        for (int b = R - 1; k < RB && b >= 0; --b) regs[k++] = b;
        for (int i = 0; i < RB; ++i)
            for (int j = i + 1; j < RB; ++j)
                if (regs[j] < regs[i]) {
                    int t = regs[i];
                    regs[i] = regs[j];
                    regs[j] = t;
                }
        pl.ph[np++] = make_phase(regs, RB, act, TB);
        done |= act;
    }
    pl.NP = np;
    return pl;
}

// register-index -> local-bit pattern (compile-time per phase and j)
constexpr unsigned reg_pattern(const Phase& p, int RB, int j)
{
    unsigned u = 0;
    for (int b = 0; b < RB; ++b)
        if ((j >> b) & 1) u |= 1u << p.reg[b];
    return u;
}

}  // namespace bk
