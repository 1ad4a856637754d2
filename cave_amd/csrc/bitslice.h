// bitslice.h -- bit-sliced cell arithmetic shared by the streaming and the
// batched kernels.  The layout and mask helpers are host + device; the adders
// and the rule are built on gfx950's v_bitop3 and are device code.
//
// A "plane word" holds one bit of 32 cells.  Cell states (4 bits) become four
// plane words P0..P3; a live mask is one plane word; neighbour counts are
// small binary numbers held as 2..5 plane words.  Every operation below is a
// handful of bitwise instructions per 32 cells.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#define CAVE_HD __host__ __device__ __forceinline__

namespace cave {

template <int N>
struct Bits {
    static constexpr int n = N;
    uint32_t b[N];
};

// A count that is always zero (a neighbour class the rule does not use).
template <>
struct Bits<0> {
    static constexpr int n = 0;
};

// ---- nibble words <-> bit planes ----
// Eight independent 4x4 bit transposes (one per nibble position i): word j,
// bit 4i+k  <->  plane k, bit 4i+j.  The transform is its own inverse.  For a
// group of four consecutive packed words (32 cells of one x-row) the planes
// use the "interleaved" cell order: bit 4i+j holds cell 8j+i of the group.
CAVE_HD void transpose4x4(uint32_t &w0, uint32_t &w1, uint32_t &w2, uint32_t &w3)
{
    // swap the off-diagonal 2x2 sub-blocks (row pairs 0/2 and 1/3, shift 2)
    uint32_t a0 = (w0 & 0x33333333u) | ((w2 << 2) & 0xCCCCCCCCu);
    uint32_t a2 = ((w0 >> 2) & 0x33333333u) | (w2 & 0xCCCCCCCCu);
    uint32_t a1 = (w1 & 0x33333333u) | ((w3 << 2) & 0xCCCCCCCCu);
    uint32_t a3 = ((w1 >> 2) & 0x33333333u) | (w3 & 0xCCCCCCCCu);
    // swap inside each 2x2 block (row pairs 0/1 and 2/3, shift 1)
    w0 = (a0 & 0x55555555u) | ((a1 << 1) & 0xAAAAAAAAu);
    w1 = ((a0 >> 1) & 0x55555555u) | (a1 & 0xAAAAAAAAu);
    w2 = (a2 & 0x55555555u) | ((a3 << 1) & 0xAAAAAAAAu);
    w3 = ((a2 >> 1) & 0x55555555u) | (a3 & 0xAAAAAAAAu);
}

// ---- live / non-zero masks ----
// nm[k] = ~0 where bit k of M is 0, else 0; en = ~0 unless M = 16 (no
// representable live state, searchSimulationCube.slang:141 with 4-bit cells).
CAVE_HD uint32_t live_mask(const uint32_t P[4], const uint32_t nm[4], uint32_t en)
{
    return en & (P[0] ^ nm[0]) & (P[1] ^ nm[1]) & (P[2] ^ nm[2]) & (P[3] ^ nm[3]);
}

CAVE_HD uint32_t nonzero_mask(const uint32_t P[4]) { return P[0] | P[1] | P[2] | P[3]; }

// ---- x-neighbour live masks ----
// Interleaved order (bit 4i+j = cell 8j+i).  prev/next are the live masks of
// the groups at x-32 / x+32 (0 outside the grid).
CAVE_HD uint32_t shift_left_neighbor_il(uint32_t a, uint32_t prev)
{   // bit b <- live(cell(b) - 1)
    return (a << 4) | ((a >> 27) & 0xEu) | (prev >> 31);
}
CAVE_HD uint32_t shift_right_neighbor_il(uint32_t a, uint32_t next)
{   // bit b <- live(cell(b) + 1)
    return (a >> 4) | ((a & 0xEu) << 27) | (next << 31);
}
// Natural order (bit c = cell c).
CAVE_HD uint32_t shift_left_neighbor_nat(uint32_t a, uint32_t prev) { return (a << 1) | (prev >> 31); }
CAVE_HD uint32_t shift_right_neighbor_nat(uint32_t a, uint32_t next) { return (a >> 1) | (next << 31); }

// ---- gfx950 three-input bit operations ----
// v_bitop3_b32 evaluates any function of three words in one instruction; its
// truth table is the function applied to S0 = 0xF0, S1 = 0xCC, S2 = 0xAA.
__device__ __forceinline__ uint32_t xor3(uint32_t a, uint32_t b, uint32_t c)
{   // a ^ b ^ c
    return __builtin_amdgcn_bitop3_b32(a, b, c, (unsigned char)0x96);
}

__device__ __forceinline__ uint32_t maj3(uint32_t a, uint32_t b, uint32_t c)
{   // majority(a, b, c): the carry of a + b + c
    return __builtin_amdgcn_bitop3_b32(a, b, c, (unsigned char)0xE8);
}

__device__ __forceinline__ uint32_t xor_and(uint32_t a, uint32_t b, uint32_t c)
{   // (a ^ b) & c
    return __builtin_amdgcn_bitop3_b32(a, b, c, (unsigned char)0x28);
}

__device__ __forceinline__ uint32_t and_or(uint32_t a, uint32_t b, uint32_t c)
{   // (a & b) | c
    return __builtin_amdgcn_bitop3_b32(a, b, c, (unsigned char)0xEA);
}

// ---- bit-sliced adders ----
// The compiler cannot simplify through v_bitop3, so the adders decide at
// compile time which words of each column exist (planes of the operands,
// carries of the previous column) and add only those: three words are one
// xor3 + one maj3, two words one XOR + one AND, one word passes through.

// sum and carry of the first P words of x
template <int P>
__device__ __forceinline__ void column_add(const uint32_t *x, uint32_t &s, uint32_t &c)
{
    if constexpr (P == 3) {
        s = xor3(x[0], x[1], x[2]);
        c = maj3(x[0], x[1], x[2]);
    } else if constexpr (P == 2) {
        s = x[0] ^ x[1];
        c = x[0] & x[1];
    } else if constexpr (P == 1) {
        s = x[0];
        c = 0u;
    } else {
        s = 0u;
        c = 0u;
    }
}

// columns I.. of a + b + c; k holds the NK carries out of column I - 1
template <int W, int I, int NK, int NA, int NB, int NC>
__device__ __forceinline__ void add3_columns(const Bits<NA> &a, const Bits<NB> &b, const Bits<NC> &c,
                                             const uint32_t *k, Bits<W> &r)
{
    if constexpr (I < W) {
        constexpr int P = (I < NA) + (I < NB) + (I < NC);
        uint32_t x[3] = {0u, 0u, 0u};
        int n = 0;
        if constexpr (I < NA) x[n++] = a.b[I];
        if constexpr (I < NB) x[n++] = b.b[I];
        if constexpr (I < NC) x[n++] = c.b[I];
        uint32_t s1, c1;
        column_add<P>(x, s1, c1);
        constexpr int P2 = (P > 0) + NK;
        uint32_t y[3] = {0u, 0u, 0u};
        int m = 0;
        if constexpr (P > 0) y[m++] = s1;
        if constexpr (NK >= 1) y[m++] = k[0];
        if constexpr (NK >= 2) y[m++] = k[1];
        uint32_t s2, c2;
        column_add<P2>(y, s2, c2);
        r.b[I] = s2;
        uint32_t kn[2] = {0u, 0u};
        int q = 0;
        if constexpr (P >= 2) kn[q++] = c1;
        if constexpr (P2 >= 2) kn[q++] = c2;
        add3_columns<W, I + 1, (P >= 2) + (P2 >= 2)>(a, b, c, kn, r);
    }
}

// a + b + c, W output planes (the caller guarantees the sum fits).
template <int W, int NA, int NB, int NC>
__device__ __forceinline__ Bits<W> add3(const Bits<NA> &a, const Bits<NB> &b, const Bits<NC> &c)
{
    Bits<W> r;
    const uint32_t k[2] = {0u, 0u};
    add3_columns<W, 0, 0>(a, b, c, k, r);
    return r;
}

// columns I.. of a + b; K: column I - 1 carries out
template <int W, int I, bool K, int NA, int NB>
__device__ __forceinline__ void add2_columns(const Bits<NA> &a, const Bits<NB> &b, uint32_t k, Bits<W> &r)
{
    if constexpr (I < W) {
        constexpr int P = (I < NA) + (I < NB) + (K ? 1 : 0);
        uint32_t x[3] = {0u, 0u, 0u};
        int n = 0;
        if constexpr (I < NA) x[n++] = a.b[I];
        if constexpr (I < NB) x[n++] = b.b[I];
        if constexpr (K) x[n++] = k;
        uint32_t s, c;
        column_add<P>(x, s, c);
        r.b[I] = s;
        add2_columns<W, I + 1, (P >= 2)>(a, b, c, r);
    }
}

// a + b, W output planes, as a ripple-carry adder.
template <int W, int NA, int NB>
__device__ __forceinline__ Bits<W> add2(const Bits<NA> &a, const Bits<NB> &b)
{
    Bits<W> r;
    add2_columns<W, 0, false>(a, b, 0u, r);
    return r;
}

// ---- per-row x-sums ----
// From a row's live mask A and its x-neighbour masks L/R:
//   A  : the cell itself          (1 plane)
//   LR : L + R                    (2 planes, <= 2)
//   XS : A + L + R                (2 planes, <= 3)
struct RowSums {
    uint32_t a;
    Bits<2> lr;
    Bits<2> xs;
};

__device__ __forceinline__ RowSums row_sums(uint32_t a, uint32_t l, uint32_t r)
{   // XS is one full adder; LR is dropped by the compiler where no kind uses it
    RowSums s;
    s.a = a;
    s.lr.b[0] = l ^ r;
    s.lr.b[1] = l & r;
    s.xs.b[0] = xor3(a, l, r);
    s.xs.b[1] = maj3(a, l, r);
    return s;
}

// Contribution kinds of the 3x3 (y, z) rows around a cell, by the L1 distance
// class of the row offset (|dy|+|dz|): the row's own cell has class c0 and its
// x-neighbours class c0+1.  With F (class 1), E (class 2) and C (class 3):
//   k0 (centre row, c0 = 0) = F*(L+R)
//   k1 (face rows,  c0 = 1) = F*A + E*(L+R)
//   k2 (edge rows,  c0 = 2) = E*A + C*(L+R)
// Then, per cell, with S0(z) = k1(y-1) + k0(y) + k1(y+1) and
// S1(z) = k2(y-1) + k1(y) + k2(y+1) at plane z:
//   N = S1(z-1) + S0(z) + S1(z+1).
// Each kind has only the planes it can use: Bits<2> for the two-neighbour
// sums, Bits<1> for the cell alone, Bits<0> when the classes are off.
template <uint32_t FLAGS>
CAVE_HD auto kind0(const RowSums &s)
{
    if constexpr ((FLAGS & 1u) != 0u) return s.lr;
    else return Bits<0>{};
}

template <uint32_t FLAGS>
CAVE_HD auto kind1(const RowSums &s)
{
    constexpr bool F = FLAGS & 1u, E = FLAGS & 2u;
    if constexpr (F && E) return s.xs;
    else if constexpr (E) return s.lr;
    else if constexpr (F) return Bits<1>{{s.a}};
    else return Bits<0>{};
}

template <uint32_t FLAGS>
CAVE_HD auto kind2(const RowSums &s)
{
    constexpr bool E = FLAGS & 2u, C = FLAGS & 4u;
    if constexpr (E && C) return s.xs;
    else if constexpr (C) return s.lr;
    else if constexpr (E) return Bits<1>{{s.a}};
    else return Bits<0>{};
}

// ---- rule ----

// Evaluates the born / survive tables for 32 cells at once: for every count v
// that appears in either table, the minterm [N == v] is built from the five
// count planes and OR-ed into the matching result.  `tables` = born | surv.
CAVE_HD void apply_rule(const Bits<5> &n, uint32_t born_lut, uint32_t surv_lut, uint32_t &tb, uint32_t &ts)
{
    tb = 0u;
    ts = 0u;
    uint32_t u = born_lut | surv_lut;
    while (u) {
        const uint32_t v = (uint32_t)__builtin_ctz(u);
        u &= u - 1u;
        uint32_t mt = ~0u;
#pragma unroll
        for (int b = 0; b < 5; b++) mt &= n.b[b] ^ (((v >> b) & 1u) ? 0u : ~0u);
        tb |= mt & (0u - ((born_lut >> v) & 1u));
        ts |= mt & (0u - ((surv_lut >> v) & 1u));
    }
}

// live_mask as four (a ^ b) & c steps, with `en` folded into the first.
__device__ __forceinline__ uint32_t live_mask_v3(const uint32_t P[4], const uint32_t nm[4], uint32_t en)
{
    uint32_t m = xor_and(P[3], nm[3], en);
    m = xor_and(P[2], nm[2], m);
    m = xor_and(P[1], nm[1], m);
    return xor_and(P[0], nm[0], m);
}

__device__ __forceinline__ uint32_t xnor_and(uint32_t a, uint32_t b, uint32_t c)
{   // ~(a ^ b) & c
    return __builtin_amdgcn_bitop3_b32(a, b, c, (unsigned char)0x82);
}

// Same tables for R rows at once.  The scalar loop over the counts present in
// either table (a wavefront-uniform value) is shared by the rows; per count
// and row the minterm [N == v] is one XNOR and four XNOR-ANDs against the
// count's bits, and each table takes it with one AND-OR against a uniform
// all-ones / zero mask -- seven instructions, no branches.
template <int R>
__device__ __forceinline__ void apply_rule_rows(const Bits<5> *n, uint32_t born_lut, uint32_t surv_lut, uint32_t *tb,
                                                uint32_t *ts)
{
#pragma unroll
    for (int r = 0; r < R; r++) tb[r] = ts[r] = 0u;
    uint32_t u = born_lut | surv_lut;
    while (u) {
        const uint32_t v = (uint32_t)__builtin_ctz(u);
        u &= u - 1u;
        const uint32_t bm = 0u - ((born_lut >> v) & 1u);
        const uint32_t sm = 0u - ((surv_lut >> v) & 1u);
        // pol[b] = ~0 where bit b of the count is 1: a one-bit sign-extending
        // field extract, one scalar op per bit.  readfirstlane keeps each mask
        // in a scalar register (the count is the same in every lane): left to
        // itself the compiler builds the top bit's mask as a vector select
        // (one VALU op per count).
        uint32_t pol[5];
#pragma unroll
        for (int b = 0; b < 5; b++)
            pol[b] = __builtin_amdgcn_readfirstlane((uint32_t)__builtin_amdgcn_sbfe((int)v, b, 1));
#pragma unroll
        for (int r = 0; r < R; r++) {
            uint32_t mt = ~(n[r].b[0] ^ pol[0]);
#pragma unroll
            for (int b = 1; b < 5; b++) mt = xnor_and(n[r].b[b], pol[b], mt);
            tb[r] = and_or(mt, bm, tb[r]);
            ts[r] = and_or(mt, sm, ts[r]);
        }
    }
}

// ---- state update ----
__device__ __forceinline__ uint32_t and_not(uint32_t a, uint32_t b)
{   // a & ~b
    return __builtin_amdgcn_bitop3_b32(a, b, b, (unsigned char)0x30);
}

// new = M where set_m, s-1 where s != 0 and !set_m, 0 elsewhere
// (searchSimulationCube.slang:178-211).  mk[k] = ~0 where bit k of M is set.
// The decrement is built already masked by keep = nz & ~set_m: bit k of
// s - 1 is P[k] XOR the borrow (~P[0] & .. & ~P[k-1]), so each masked plane
// keep & (P[k] ^ borrow) is one three-input op.
__device__ __forceinline__ void update_planes(uint32_t P[4], uint32_t set_m, uint32_t nz, const uint32_t mk[4])
{
    const uint32_t keep = and_not(nz, set_m);
    const uint32_t e0 = and_not(keep, P[0]);
    const uint32_t e1 = __builtin_amdgcn_bitop3_b32(keep, P[1], P[0], (unsigned char)0x90);  // keep & ~(P1 ^ P0)
    uint32_t bor = __builtin_amdgcn_bitop3_b32(P[0], P[1], P[1], (unsigned char)0x03);     // ~(P0 | P1)
    const uint32_t e2 = __builtin_amdgcn_bitop3_b32(keep, P[2], bor, (unsigned char)0x60);  // keep & (P2 ^ bor)
    bor = and_not(bor, P[2]);
    const uint32_t e3 = __builtin_amdgcn_bitop3_b32(keep, P[3], bor, (unsigned char)0x60);
    P[0] = and_or(set_m, mk[0], e0);
    P[1] = and_or(set_m, mk[1], e1);
    P[2] = and_or(set_m, mk[2], e2);
    P[3] = and_or(set_m, mk[3], e3);
}

// Rules whose tables hold at most kRuleSlots counts in all are evaluated from
// masks prepared once per rule (RuleConsts::slot_*), with no scalar work per
// step; larger rules take the count loop of apply_rule_rows.
constexpr int kRuleSlots = 4;

// Per-rule constants in the form the kernels use.
struct RuleConsts {
    uint32_t nm[4];   // ~0 where bit k of M is 0
    uint32_t mk[4];   // ~0 where bit k of (M & 15) is 1
    uint32_t en;      // ~0 unless M == 16
    uint32_t born_lut, surv_lut;
    // per slot: ~0 where bit b of the slot's count is 1, and whether the count
    // is in the born / survive table; unused slots hold count 31, which no
    // neighbour count reaches (at most 27)
    uint32_t slot_pol[kRuleSlots][5];
    uint32_t slot_b[kRuleSlots], slot_s[kRuleSlots];
    uint32_t use_slots;  // 1: the tables hold at most kRuleSlots counts
};

CAVE_HD RuleConsts rule_consts(uint32_t M, uint32_t born_lut, uint32_t surv_lut)
{
    RuleConsts c;
    for (int k = 0; k < 4; k++) {
        c.nm[k] = ((M >> k) & 1u) ? 0u : ~0u;
        c.mk[k] = ((M >> k) & 1u) ? ~0u : 0u;
    }
    c.en = M <= 15u ? ~0u : 0u;
    c.born_lut = born_lut;
    c.surv_lut = surv_lut;
    uint32_t u = born_lut | surv_lut;
    c.use_slots = __builtin_popcount(u) <= kRuleSlots ? 1u : 0u;
    for (int k = 0; k < kRuleSlots; k++) {
        uint32_t v = 31u, b = 0u, s = 0u;
        if (u) {
            v = (uint32_t)__builtin_ctz(u);
            u &= u - 1u;
            b = ((born_lut >> v) & 1u) ? ~0u : 0u;
            s = ((surv_lut >> v) & 1u) ? ~0u : 0u;
        }
        for (int i = 0; i < 5; i++) c.slot_pol[k][i] = ((v >> i) & 1u) ? ~0u : 0u;
        c.slot_b[k] = b;
        c.slot_s[k] = s;
    }
    return c;
}

// The born / survive tables for R rows: from the prepared slots when the rule
// has at most kRuleSlots counts (seven instructions per slot and row, no
// scalar work), else by the count loop.
template <int R>
__device__ __forceinline__ void apply_rule_consts(const Bits<5> *n, const RuleConsts &rc, uint32_t *tb, uint32_t *ts)
{
    if (rc.use_slots) {
#pragma unroll
        for (int r = 0; r < R; r++) tb[r] = ts[r] = 0u;
#pragma unroll
        for (int k = 0; k < kRuleSlots; k++) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                uint32_t mt = ~(n[r].b[0] ^ rc.slot_pol[k][0]);
#pragma unroll
                for (int b = 1; b < 5; b++) mt = xnor_and(n[r].b[b], rc.slot_pol[k][b], mt);
                tb[r] = and_or(mt, rc.slot_b[k], tb[r]);
                ts[r] = and_or(mt, rc.slot_s[k], ts[r]);
            }
        }
    } else {
        apply_rule_rows<R>(n, rc.born_lut, rc.surv_lut, tb, ts);
    }
}

// FEV (all 26 neighbours): the count is the 3x3x3 box sum B of live cells
// minus the cell itself, and B = S1(z-1) + S1(z) + S1(z+1) (with F, E and C
// all set, k0 = k1 - A and k2 = k1, so S0 = S1 - A at the centre row).  The
// kernels then sum S1 only and apply the rule to B: birth applies to dead
// cells (not live, so N = B) and survival to live cells (N = B - 1), i.e. the
// survival table shifted up by one count (B <= 27 fits the 32-bit table).
template <uint32_t FLAGS>
inline constexpr bool kBoxSum = (FLAGS & 7u) == 7u;

CAVE_HD RuleConsts rule_consts_for(uint32_t flags, uint32_t M, uint32_t born_lut, uint32_t surv_lut)
{
    return rule_consts(M, born_lut, (flags & 7u) == 7u ? surv_lut << 1 : surv_lut);
}

}  // namespace cave
