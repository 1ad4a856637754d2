// bitslice.h -- bit-sliced cell arithmetic shared by the streaming and the
// batched kernels (host + device, so the pieces are unit-testable on a CPU).
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
    uint32_t b[N];
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

// ---- bit-sliced adders ----
CAVE_HD void full_add(uint32_t x, uint32_t y, uint32_t z, uint32_t &s, uint32_t &c)
{
    const uint32_t t = x ^ y;
    s = t ^ z;
    c = (t & z) | (~t & x);
}

// a + b + c, W output planes (the caller guarantees the sum fits).  Column i
// receives a_i, b_i, c_i and up to two carries of column i-1; constant-zero
// planes fold away at compile time.
template <int W, int NA, int NB, int NC>
CAVE_HD Bits<W> add3(const Bits<NA> &a, const Bits<NB> &b, const Bits<NC> &c)
{
    Bits<W> r;
    uint32_t k1 = 0, k2 = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint32_t ai = i < NA ? a.b[i < NA ? i : 0] : 0u;
        const uint32_t bi = i < NB ? b.b[i < NB ? i : 0] : 0u;
        const uint32_t ci = i < NC ? c.b[i < NC ? i : 0] : 0u;
        uint32_t s1, c1, s2, c2;
        full_add(ai, bi, ci, s1, c1);
        full_add(s1, k1, k2, s2, c2);
        r.b[i] = s2;
        k1 = c1;
        k2 = c2;
    }
    return r;
}

// a + b, W output planes, as a ripple-carry adder with one carry per column
// (sum = a ^ b ^ k, carry = majority(a, b, k): one three-input bit operation
// each).  add3 with a zero third operand costs two half adders per column.
template <int W, int NA, int NB>
CAVE_HD Bits<W> add2(const Bits<NA> &a, const Bits<NB> &b)
{
    Bits<W> r;
    uint32_t k = 0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        const uint32_t ai = i < NA ? a.b[i < NA ? i : 0] : 0u;
        const uint32_t bi = i < NB ? b.b[i < NB ? i : 0] : 0u;
        r.b[i] = ai ^ bi ^ k;
        k = (ai & bi) | (k & (ai ^ bi));
    }
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

CAVE_HD RowSums row_sums(uint32_t a, uint32_t l, uint32_t r)
{
    RowSums s;
    s.a = a;
    s.lr.b[0] = l ^ r;
    s.lr.b[1] = l & r;
    s.xs.b[0] = a ^ s.lr.b[0];
    s.xs.b[1] = s.lr.b[1] | (a & s.lr.b[0]);
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
template <uint32_t FLAGS>
CAVE_HD Bits<2> kind0(const RowSums &s)
{
    Bits<2> k{{0u, 0u}};
    if (FLAGS & 1u) k = s.lr;
    return k;
}

template <uint32_t FLAGS>
CAVE_HD Bits<2> kind1(const RowSums &s)
{
    constexpr bool F = FLAGS & 1u, E = FLAGS & 2u;
    if (F && E) return s.xs;
    if (E) return s.lr;
    Bits<2> k{{F ? s.a : 0u, 0u}};
    return k;
}

template <uint32_t FLAGS>
CAVE_HD Bits<2> kind2(const RowSums &s)
{
    constexpr bool E = FLAGS & 2u, C = FLAGS & 4u;
    if (E && C) return s.xs;
    if (C) return s.lr;
    Bits<2> k{{E ? s.a : 0u, 0u}};
    return k;
}

// ---- rule ----
// gfx950 three-input bit operations (v_bitop3_b32; truth table over S0 = 0xF0,
// S1 = 0xCC, S2 = 0xAA): one instruction for a XOR feeding an AND, and for an
// AND feeding an OR.
__device__ __forceinline__ uint32_t xor_and(uint32_t a, uint32_t b, uint32_t c)
{   // (a ^ b) & c
    return __builtin_amdgcn_bitop3_b32(a, b, c, (unsigned char)0x28);
}

__device__ __forceinline__ uint32_t and_or(uint32_t a, uint32_t b, uint32_t c)
{   // (a & b) | c
    return __builtin_amdgcn_bitop3_b32(a, b, c, (unsigned char)0xEA);
}

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

// Same tables for R rows at once.  The scalar loop over the counts present in
// either table (a wavefront-uniform value) is shared by the rows; per count
// and row the minterm [N == v] is a chain of five XOR-ANDs (the XOR applies
// the count bit's polarity), and each table takes it with one AND-OR against
// a uniform all-ones / zero mask -- seven instructions, no branches.
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
        // inv[b] = ~0 where bit b of the count is 0.  The empty asm pins each
        // mask to a scalar register: left to itself the compiler builds the
        // top bit's mask as a vector select (one VALU op per count).
        uint32_t inv[5];
#pragma unroll
        for (int b = 0; b < 5; b++) {
            inv[b] = ((v >> b) & 1u) - 1u;
            asm volatile("" : "+s"(inv[b]));
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            uint32_t mt = n[r].b[0] ^ inv[0];
#pragma unroll
            for (int b = 1; b < 5; b++) mt = xor_and(n[r].b[b], inv[b], mt);
            tb[r] = and_or(mt, bm, tb[r]);
            ts[r] = and_or(mt, sm, ts[r]);
        }
    }
}

// ---- state update ----
// new = M where set_m, s-1 where s != 0 and !set_m, 0 elsewhere
// (searchSimulationCube.slang:178-211).  mk[k] = ~0 where bit k of M is set.
CAVE_HD void update_planes(uint32_t P[4], uint32_t set_m, uint32_t nz, const uint32_t mk[4])
{
    uint32_t bor = ~P[0];
    uint32_t d0 = ~P[0];
    uint32_t d1 = P[1] ^ bor;
    bor &= ~P[1];
    uint32_t d2 = P[2] ^ bor;
    bor &= ~P[2];
    uint32_t d3 = P[3] ^ bor;
    const uint32_t keep = nz & ~set_m;
    P[0] = (set_m & mk[0]) | (keep & d0);
    P[1] = (set_m & mk[1]) | (keep & d1);
    P[2] = (set_m & mk[2]) | (keep & d2);
    P[3] = (set_m & mk[3]) | (keep & d3);
}

// Per-rule constants in the form the kernels use.
struct RuleConsts {
    uint32_t nm[4];   // ~0 where bit k of M is 0
    uint32_t mk[4];   // ~0 where bit k of (M & 15) is 1
    uint32_t en;      // ~0 unless M == 16
    uint32_t born_lut, surv_lut;
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
    return c;
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
