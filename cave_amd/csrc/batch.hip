// batch.hip -- persistent, work-stealing batched simulation kernel.
//
// One launch evaluates a whole list of simulations (rule config x grid x
// seed).  Each block is a persistent worker: it takes the next job index from
// a global atomic counter (the list is sorted by decreasing cost on the host,
// so the queue hands out the longest jobs first), builds the initial cells in
// place from the cell hash, runs the job's ticks with the search manager
// fused in (searchManager.slang:98-149: totals, max / final alive, early death,
// tick budget), writes the result and takes the next job.  Nothing returns to
// the host between ticks or between jobs.
//
// Cell storage per job (W = ceil(X/32) words per row, natural bit order):
//   * 4 state bit planes, interleaved: the four plane words of cell word
//     s = w + W*(y + Y*z) are words 4s .. 4s+3 (one 16-byte access), updated
//     in place;
//   * 2 live planes (ping-pong) with a zero guard band: rows y = -1 and Y and
//     planes z = -1 and Z exist and stay 0, each row ends with a zero guard
//     word, and one zero word precedes the plane: word 1 + w + (W+1)*((y+1) +
//     (Y+2)*(z+1)) of the plane.  The x-neighbour words of a row's first and
//     last words are guard words, so neighbour reads never branch or mask on
//     the grid faces.
// A job whose storage fits the block's LDS keeps its whole grid in LDS for all
// of its ticks ("tier 0"); a larger job runs in a per-block region of a global
// arena ("tier 1", L2 / MALL resident at search grid sizes).  Both tiers run
// the same bit-sliced tick code, instantiated per memory space.
//
// Per tick, a thread owns "items": an x-word column (w, y) over a z-chunk.  It
// walks z with the plane partial sums in registers (the separable counting of
// the streaming step kernel), reading the live words of the 3 x 3 neighbour
// rows and writing the new state of its own rows.
#include "batch.h"

namespace cave {

struct JobGeo {
    uint32_t X, Y, Z, W;
    uint32_t pplane;     // W * Y * Z: cell words of one state plane
    uint32_t aplane;     // (W+1) * (Y+2) * (Z+2) + 1: one guarded live plane
    uint32_t aslab;      // (W+1) * (Y+2): words per z plane of a live plane
    uint32_t alive0;     // first row word (w = 0, y = -1, z = -1) of live plane 0
    uint32_t last_mask;  // valid cells of the last word of a row
    uint32_t tzc;        // planes per z-chunk
    uint32_t zc0, nzc;   // this part's z-chunks [zc0, zc0 + nzc)
    uint32_t zlo, zhi;   // this part's planes [zlo, zhi)
    uint32_t nitems;     // W * Y * nzc
};

__device__ __forceinline__ JobGeo job_geo(const DevJob &j)
{
    JobGeo g;
    g.X = j.X;
    g.Y = j.Y;
    g.Z = j.Z;
    g.W = (j.X + 31u) / 32u;
    g.pplane = g.W * j.Y * j.Z;
    g.aslab = (g.W + 1u) * (j.Y + 2u);
    g.aplane = g.aslab * (j.Z + 2u) + 1u;
    g.alive0 = 4u * g.pplane + 1u;
    const uint32_t r = j.X - 32u * (g.W - 1u);
    g.last_mask = r >= 32u ? ~0u : ((1u << r) - 1u);
    const uint32_t zchunks = max(1u, j.zchunks);
    const uint32_t parts = max(1u, j.parts);
    g.tzc = (j.Z + zchunks - 1u) / zchunks;
    const uint32_t used = (j.Z + g.tzc - 1u) / g.tzc;  // non-empty chunks
    const uint32_t c = zchunks / parts;
    g.zc0 = min(j.part * c, used);
    g.nzc = min(g.zc0 + c, used) - g.zc0;
    g.zlo = min(g.zc0 * g.tzc, j.Z);
    g.zhi = min((g.zc0 + g.nzc) * g.tzc, j.Z);
    g.nitems = g.W * j.Y * g.nzc;
    return g;
}

struct BCarry {
    Bits<5> Q;    // S1 of the plane before the last input plane + S0 of the last
    Bits<4> S1p;  // S1 of the last input plane
    uint32_t A;   // own live word of the last input plane
};

// Live-plane words that another block reads or writes (the first and last
// plane of a part's slab, and the planes just outside it) go through
// device-scope atomic loads / stores: on gfx950 these are `sc1` accesses that
// are coherent across the XCDs' L2 caches without a cache-wide write-back or
// invalidate, so a split grid does not disturb the other blocks' cached cells.
template <bool COH, typename T>
__device__ __forceinline__ uint32_t ld_live(const T *p, bool shared)
{
    if constexpr (COH) {
        if (shared) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return *p;
}

template <bool COH, typename T>
__device__ __forceinline__ void st_live(T *p, uint32_t v, bool shared)
{
    if constexpr (COH) {
        if (shared) {
            __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
    }
    *p = v;
}

// A job's storage words, addressed by 32-bit word offsets.
//   * Words<false>: the block's LDS (a pointer the compiler resolves to LDS,
//     so accesses are ds_* instructions with 32-bit addresses);
//   * Words<true>: global memory (the block's arena region or the job's own
//     region) as an address_space(1) base plus a 32-bit byte offset, which
//     the compiler turns into global accesses with the base in scalar
//     registers instead of 64-bit address arithmetic per access.  Regions
//     stay far below 2^30 words, so the byte offset fits 32 bits.
typedef __attribute__((address_space(1))) uint32_t gword;
typedef __attribute__((address_space(1))) uint8_t gbyte;

// ld4 / st4: the four state planes of one cell word, 16 bytes at a word
// offset that is a multiple of 4 (every storage base is 16-byte aligned).
template <bool GLOBAL>
struct Words {
    uint32_t *base;
    __device__ __forceinline__ uint32_t *at(uint32_t off) const { return base + off; }
    __device__ __forceinline__ uint4 ld4(uint32_t off) const { return *reinterpret_cast<const uint4 *>(base + off); }
    __device__ __forceinline__ void st4(uint32_t off, const uint4 &v) const
    {
        *reinterpret_cast<uint4 *>(base + off) = v;
    }
};

template <>
struct Words<true> {
    typedef unsigned int quad __attribute__((ext_vector_type(4)));
    typedef __attribute__((address_space(1))) quad gquad;
    gword *base;
    __device__ __forceinline__ gword *at(uint32_t off) const
    {
        return reinterpret_cast<gword *>(reinterpret_cast<gbyte *>(base) + (off << 2));
    }
    __device__ __forceinline__ uint4 ld4(uint32_t off) const
    {
        const quad q = *reinterpret_cast<const gquad *>(reinterpret_cast<gbyte *>(base) + (off << 2));
        return make_uint4(q.x, q.y, q.z, q.w);
    }
    __device__ __forceinline__ void st4(uint32_t off, const uint4 &v) const
    {
        const quad q = {v.x, v.y, v.z, v.w};
        *reinterpret_cast<gquad *>(reinterpret_cast<gbyte *>(base) + (off << 2)) = q;
    }
};

// One item: column (w, y), planes [z0, z1).  acur / anext are the first row
// words of the current / next live plane.  The x-neighbour words of a row
// are the previous / next words, guard words (zero) at the row's ends.
// COH: the job is split over several blocks (boundary planes are shared).
template <uint32_t FLAGS, bool COH, bool GLOBAL>
__device__ __forceinline__ void run_item(const JobGeo &g, const Words<GLOBAL> &S, uint32_t acur, uint32_t anext,
                                         const RuleConsts &rc, int w, int y, int z0, int z1, StatAcc &acc)
{
    const uint32_t vmask = (uint32_t)w + 1u < g.W ? ~0u : g.last_mask;
    uint32_t colmask = 0u;
    int zmin = 0x7fffffff, zmax = -1;
    // Running word offsets for the step at plane zi (the first step has
    // zi = z0 - 1), advanced by one plane per step instead of recomputed with
    // multiplies:
    //   aoff: live word of row y-1 at plane zi:      acur + w + (W+1)*(y + (Y+2)*(zi+1))
    //   poff: state words of row y at plane zi-1:    4 * (w + W*(y + Y*(zi-1)))
    //   noff: next live word of row y at plane zi-1: anext + w + (W+1)*((y+1) + (Y+2)*zi)
    // poff and noff wrap below plane 0 on the first steps; they are used only
    // once plane zi-1 is one of the item's planes.
    const uint32_t wy = (uint32_t)w + g.W * (uint32_t)y;
    const uint32_t pstride = g.W * g.Y;
    const uint32_t arow = g.W + 1u;  // live row stride
    const uint32_t wya = (uint32_t)w + arow * (uint32_t)y;
    uint32_t aoff = acur + wya + g.aslab * (uint32_t)z0;
    uint32_t poff = 4u * (wy + pstride * (uint32_t)(z0 - 2));
    uint32_t noff = anext + wya + arow + g.aslab * (uint32_t)(z0 - 1);

    auto step = [&](int zi, const BCarry &ci, BCarry &co) {
        RowSums rs[3];
        // planes zlo-1 and zhi are written by the neighbouring parts: read
        // coherently.  This part's own planes, zlo and zhi-1 included, were
        // written by this block, whose writes it sees after a barrier.
        const bool shr = COH && (zi == (int)g.zlo - 1 || zi == (int)g.zhi);
        uint32_t av[3], pv[3], nv[3];
        if (shr) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const auto *q = S.at(aoff - 1u + (uint32_t)d * arow);  // the word before the row's own
                pv[d] = ld_live<COH>(q, true);
                av[d] = ld_live<COH>(q + 1, true);
                nv[d] = ld_live<COH>(q + 2, true);
            }
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                // one address per row: the word before the row's own, then +1, +2
                const auto *q = S.at(aoff - 1u + (uint32_t)d * arow);
                pv[d] = q[0];
                av[d] = q[1];
                nv[d] = q[2];
            }
        }
#pragma unroll
        for (int d = 0; d < 3; d++)
            rs[d] = row_sums(av[d], shift_left_neighbor_nat(av[d], pv[d]), shift_right_neighbor_nat(av[d], nv[d]));
        const Bits<4> S1c = add3<4>(kind2<FLAGS>(rs[0]), kind1<FLAGS>(rs[1]), kind2<FLAGS>(rs[2]));
        if constexpr (kBoxSum<FLAGS>) {
            co.Q = add2<5>(ci.S1p, S1c);
        } else {
            const Bits<4> S0c = add3<4>(kind1<FLAGS>(rs[0]), kind0<FLAGS>(rs[1]), kind1<FLAGS>(rs[2]));
            co.Q = add2<5>(ci.S1p, S0c);
        }
        co.S1p = S1c;
        co.A = rs[1].a;
        const int z = zi - 1;
        if (z >= z0) {
            const Bits<5> n = add2<5>(ci.Q, S1c);
            uint32_t tb, ts;
            apply_rule_consts<1>(&n, rc, &tb, &ts);
            const uint4 pq = S.ld4(poff);
            uint32_t pl[4] = {pq.x, pq.y, pq.z, pq.w};
            const uint32_t NZ = nonzero_mask(pl);
            const uint32_t Al = ci.A;
            const uint32_t set_m = vmask & ((Al & ts) | (~NZ & tb));
            const uint32_t surv = set_m & Al;
            acc.surv += __popc(surv);
            acc.born += __popc(set_m);  // born + surviving: the tick's caller takes surv out
            acc.dying += __popc(and_not(NZ, surv));
            update_planes(pl, set_m, NZ, rc.mk);
            S.st4(poff, make_uint4(pl[0], pl[1], pl[2], pl[3]));
            st_live<COH>(S.at(noff), set_m, COH && (z == (int)g.zlo || z + 1 == (int)g.zhi));
            colmask |= set_m;
            if (set_m) {
                zmin = min(zmin, z);
                zmax = max(zmax, z);
            }
        }
        aoff += g.aslab;
        poff += 4u * pstride;
        noff += g.aslab;
    };

    BCarry c0, c1;
    for (int k = 0; k < 5; k++) c0.Q.b[k] = 0u;
    for (int k = 0; k < 4; k++) c0.S1p.b[k] = 0u;
    c0.A = 0u;
    int zi = z0 - 1;
    for (; zi + 1 <= z1; zi += 2) {
        step(zi, c0, c1);
        step(zi + 1, c1, c0);
    }
    if (zi <= z1) step(zi, c0, c1);
    if (colmask) {
        acc.bmin[0] = min(acc.bmin[0], 32 * w + (int)__builtin_ctz(colmask));
        acc.bmax[0] = max(acc.bmax[0], 32 * w + 31 - (int)__builtin_clz(colmask));
        acc.bmin[1] = min(acc.bmin[1], y);
        acc.bmax[1] = max(acc.bmax[1], y);
        acc.bmin[2] = min(acc.bmin[2], zmin);
        acc.bmax[2] = max(acc.bmax[2], zmax);
    }
}

// ---------------------------------------------------------------------------
// The elongated rhombic dodecahedron (ERD) cell shape
// (searchSimulationERD.slang:81-97, elongatedRhombicDodecahedron.slang:6-27):
// 12 face neighbours whose offsets depend on the parity of the cell's row y,
// 4 edge neighbours, no corners.  Grouped by neighbour row (y', z'), a cell
// (x, y, z) sees:
//   row (y, z):         x-1, x+1                    (F)
//   rows (y+-1, z):     the pair P                  (F)
//   rows (y, z+-1):     x (F); x-1, x+1 (E)
//   rows (y+-1, z-1):   the pair P, when y is even  (F)
//   rows (y+-1, z+1):   the pair P, when y is odd   (F)
// where P = {x-1, x} for even y and {x, x+1} for odd y.  Per plane zi an item
// of row y therefore forms
//   T  = F*A + E*(L+R)          (row y: its part in the cells of planes zi+-1)
//   Y  = P(y-1) + P(y+1)        (F)
//   C  = F*((L+R) + Y)          (its part in the cells of plane zi)
//   Bk = T + [y even] Y         (its part in the cells of plane zi+1)
//   Fw = T + [y odd] Y          (its part in the cells of plane zi-1)
// and N(z) = Bk(z-1) + C(z) + Fw(z+1) <= 16: the cube's z-walk with two
// carried sums, the parity a per-lane mask.  Cells, guard bands and the state
// update are the cube's.
// WRAP (the W flag): every axis wraps, each offset its own term (on an axis
// shorter than 3 cells the same cell counts once per offset that reaches it).
// The loads are the same, at wrapped addresses: rows y+-1 and planes z+-1
// past a face are taken from the opposite face instead of the zero guard
// band, the first word of a row takes its left neighbour from the row's last
// word (cell X-1 shifted up to bit 31), and the last word takes the right
// neighbour of cell X-1 from bit 0 of the row's first word.
struct ECarry {
    Bits<4> Q;   // Bk of the plane before the last input plane + C of the last
    Bits<3> Bk;  // Bk of the last input plane
    uint32_t A;  // own live word of the last input plane
};

template <uint32_t FLAGS, bool COH, bool GLOBAL, bool WRAP>
__device__ __forceinline__ void run_item_erd(const JobGeo &g, const Words<GLOBAL> &S, uint32_t acur, uint32_t anext,
                                             const RuleConsts &rc, int w, int y, int z0, int z1, StatAcc &acc)
{
    constexpr bool F = (FLAGS & 1u) != 0u, E = (FLAGS & 2u) != 0u;
    const uint32_t odd = (y & 1) ? ~0u : 0u;
    // WRAP: word offsets of the x-neighbour words and of rows y-1 / y+1
    // from the row's own word, and the valid cells of a row's last word
    const bool first = w == 0, last = w == (int)g.W - 1;
    const uint32_t rlast = g.X - 32u * (g.W - 1u);  // 1..32
    uint32_t dpv = 0xFFFFFFFFu, dnv = 1u, ro[3] = {0u, g.W + 1u, 2u * (g.W + 1u)};
    if constexpr (WRAP) {
        if (first) dpv = g.W - 1u;
        if (last) dnv = 1u - g.W;
        if (y == 0) ro[0] = g.Y * (g.W + 1u);
        if (y == (int)g.Y - 1) ro[2] = 2u * (g.W + 1u) - g.Y * (g.W + 1u);
    }
    const uint32_t vmask = (uint32_t)w + 1u < g.W ? ~0u : g.last_mask;
    uint32_t colmask = 0u;
    int zmin = 0x7fffffff, zmax = -1;
    // the running word offsets of run_item
    const uint32_t wy = (uint32_t)w + g.W * (uint32_t)y;
    const uint32_t pstride = g.W * g.Y;
    const uint32_t arow = g.W + 1u;
    const uint32_t wya = (uint32_t)w + arow * (uint32_t)y;
    uint32_t aoff = acur + wya + g.aslab * (uint32_t)z0;
    uint32_t poff = 4u * (wy + pstride * (uint32_t)(z0 - 2));
    uint32_t noff = anext + wya + arow + g.aslab * (uint32_t)(z0 - 1);

    auto step = [&](int zi, const ECarry &ci, ECarry &co) {
        const bool shr = COH && (zi == (int)g.zlo - 1 || zi == (int)g.zhi);
        uint32_t av[3], pv[3], nv[3];
        if constexpr (WRAP) {
            // planes -1 and Z are planes Z-1 and 0
            const uint32_t zc = zi == -1 ? g.Z * g.aslab : (zi == (int)g.Z ? 0u - g.Z * g.aslab : 0u);
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const uint32_t own = aoff + zc + ro[d];
                pv[d] = ld_live<COH>(S.at(own + dpv), shr);
                av[d] = ld_live<COH>(S.at(own), shr);
                nv[d] = ld_live<COH>(S.at(own + dnv), shr);
            }
        } else if (shr) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const auto *q = S.at(aoff - 1u + (uint32_t)d * arow);
                pv[d] = ld_live<COH>(q, true);
                av[d] = ld_live<COH>(q + 1, true);
                nv[d] = ld_live<COH>(q + 2, true);
            }
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const auto *q = S.at(aoff - 1u + (uint32_t)d * arow);
                pv[d] = q[0];
                av[d] = q[1];
                nv[d] = q[2];
            }
        }
        uint32_t L[3], R[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            if constexpr (WRAP) {
                L[d] = shift_left_neighbor_nat(av[d], first ? pv[d] << (32u - rlast) : pv[d]);
                R[d] = last ? (av[d] >> 1) | ((nv[d] & 1u) << (rlast - 1u)) : shift_right_neighbor_nat(av[d], nv[d]);
            } else {
                L[d] = shift_left_neighbor_nat(av[d], pv[d]);
                R[d] = shift_right_neighbor_nat(av[d], nv[d]);
            }
        }
        const uint32_t a = av[1];
        auto T = [&]() {
            if constexpr (F && E) return Bits<2>{{xor3(a, L[1], R[1]), maj3(a, L[1], R[1])}};
            else if constexpr (E) return Bits<2>{{L[1] ^ R[1], L[1] & R[1]}};
            else if constexpr (F) return Bits<1>{{a}};
            else return Bits<0>{};
        }();
        Bits<3> Bk, Fw;
        if constexpr (F) {
            Bits<2> P[2];
#pragma unroll
            for (int k = 0; k < 2; k++) {
                const int d = 2 * k;  // rows y-1 and y+1
                // odd ? R : L -- the pair's other cell
                const uint32_t sel = __builtin_amdgcn_bitop3_b32(R[d], L[d], odd, (unsigned char)0xE4);
                P[k] = Bits<2>{{av[d] ^ sel, av[d] & sel}};
            }
            const Bits<3> Y = add2<3>(P[0], P[1]);
            const Bits<2> LRy{{L[1] ^ R[1], L[1] & R[1]}};
            const Bits<3> C = add2<3>(LRy, Y);
            Bits<3> Ye, Yo;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                Ye.b[k] = and_not(Y.b[k], odd);
                Yo.b[k] = Y.b[k] & odd;
            }
            Bk = add2<3>(T, Ye);
            Fw = add2<3>(T, Yo);
            co.Q = add2<4>(ci.Bk, C);
        } else {
            Bk = add2<3>(T, Bits<0>{});
            Fw = Bk;
            co.Q = add2<4>(ci.Bk, Bits<0>{});
        }
        co.Bk = Bk;
        co.A = a;
        const int z = zi - 1;
        if (z >= z0) {
            const Bits<5> n = add2<5>(ci.Q, Fw);
            uint32_t tb, ts;
            apply_rule_consts<1>(&n, rc, &tb, &ts);
            const uint4 pq = S.ld4(poff);
            uint32_t pl[4] = {pq.x, pq.y, pq.z, pq.w};
            const uint32_t NZ = nonzero_mask(pl);
            const uint32_t Al = ci.A;
            const uint32_t set_m = vmask & ((Al & ts) | (~NZ & tb));
            const uint32_t surv = set_m & Al;
            acc.surv += __popc(surv);
            acc.born += __popc(set_m);  // born + surviving: the tick's caller takes surv out
            acc.dying += __popc(and_not(NZ, surv));
            update_planes(pl, set_m, NZ, rc.mk);
            S.st4(poff, make_uint4(pl[0], pl[1], pl[2], pl[3]));
            st_live<COH>(S.at(noff), set_m, COH && (z == (int)g.zlo || z + 1 == (int)g.zhi));
            colmask |= set_m;
            if (set_m) {
                zmin = min(zmin, z);
                zmax = max(zmax, z);
            }
        }
        aoff += g.aslab;
        poff += 4u * pstride;
        noff += g.aslab;
    };

    ECarry c0, c1;
    for (int k = 0; k < 4; k++) c0.Q.b[k] = 0u;
    for (int k = 0; k < 3; k++) c0.Bk.b[k] = 0u;
    c0.A = 0u;
    int zi = z0 - 1;
    for (; zi + 1 <= z1; zi += 2) {
        step(zi, c0, c1);
        step(zi + 1, c1, c0);
    }
    if (zi <= z1) step(zi, c0, c1);
    if (colmask) {
        acc.bmin[0] = min(acc.bmin[0], 32 * w + (int)__builtin_ctz(colmask));
        acc.bmax[0] = max(acc.bmax[0], 32 * w + 31 - (int)__builtin_clz(colmask));
        acc.bmin[1] = min(acc.bmin[1], y);
        acc.bmax[1] = max(acc.bmax[1], y);
        acc.bmin[2] = min(acc.bmin[2], zmin);
        acc.bmax[2] = max(acc.bmax[2], zmax);
    }
}

// The cells a tick computes: an inclusive box of cell coordinates (empty when
// hi < lo on an axis).  Every cell outside it is 0 before the tick and has no
// live neighbour, so it stays 0 and counts as dead; see run_entry.
// Its items for this part (plan_region): word columns w0 .. w0+nw-1 of rows
// y0 .. y0+ny-1, planes [za, zb) cut into z-chunks of tz planes, n items.
// cells: the region's cells over the whole job (every part's planes, whole
// words in x), for the computed-cells count.
struct Box {
    int32_t lo[3], hi[3];
};

struct alignas(16) Region {
    uint32_t w0, nw, y0, ny, za, zb, tz, n;  // every thread reads these eight words at a tick's start
    Box box;
    uint64_t cells;
};

// ceil(a / b), b >= 1, from a float quotient corrected to the exact value:
// a few instructions, and no branch, instead of the integer division's ~35 on
// the one thread that plans the next tick.  For a < 2^22 the float product is
// within 0.75 / b of a / b (v_rcp_f32 is within 1 ulp), so its ceiling is the
// result, one less or one more, and two selects correct it; q * b stays near
// a, below 2^32.  The planner's counts (items, planes, chunks) are far below
// 2^22; larger ones take the integer division.
__device__ __forceinline__ uint32_t div_ceil_small(uint32_t a, uint32_t b)
{
    if (a >= (1u << 22)) return a / b + (a % b != 0u ? 1u : 0u);
    uint32_t q = (uint32_t)__builtin_ceilf((float)a * __builtin_amdgcn_rcpf((float)b));
    q = q * b < a ? q + 1u : q;
    q = (q > 0u && (q - 1u) * b >= a) ? q - 1u : q;
    return q;
}

// The items of a region on this part's planes: (word column, z-chunk) pairs,
// with the z-chunks cut so that the tick has about as many items as the plan
// gave the whole part (thread 0 computes it once per tick).
__device__ __forceinline__ void plan_region(Region &r, const JobGeo &g)
{
    const Box &b = r.box;
    const int za = max(b.lo[2], (int)g.zlo), zb = min(b.hi[2] + 1, (int)g.zhi);
    r.n = 0u;
    r.za = r.zb = (uint32_t)max(za, 0);
    r.w0 = r.nw = r.y0 = r.ny = 0u;
    r.tz = 1u;
    if (za >= zb || b.hi[0] < b.lo[0] || b.hi[1] < b.lo[1]) return;
    r.w0 = (uint32_t)b.lo[0] >> 5;
    r.nw = ((uint32_t)b.hi[0] >> 5) - r.w0 + 1u;
    r.y0 = (uint32_t)b.lo[1];
    r.ny = (uint32_t)(b.hi[1] - b.lo[1] + 1);
    r.za = (uint32_t)za;
    r.zb = (uint32_t)zb;
    const uint32_t nzp = r.zb - r.za, cols = r.nw * r.ny;
    const uint32_t nzc = min(nzp, max(1u, div_ceil_small(g.nitems, cols)));
    r.tz = div_ceil_small(nzp, nzc);
    r.n = cols * div_ceil_small(nzp, r.tz);
}

// a / b and a % b from rb = the float reciprocal of b (v_rcp_f32, 1 ulp).
// For a < 2^22 the float product a * rb is within 0.75 / b of a / b, so its
// truncation is the quotient, one less or one more, and one select corrects
// it (b = 1 is exact: rcp(1) = 1).  A few instructions per item, and no
// per-region constant for thread 0 to compute between ticks.
__device__ __forceinline__ uint32_t divmod_rcp(uint32_t a, uint32_t b, float rb, uint32_t &rem)
{
    const uint32_t q = (uint32_t)((float)a * rb);
    const int32_t r = (int32_t)(a - q * b), bi = (int32_t)b;
    rem = (uint32_t)(r < 0 ? r + bi : (r >= bi ? r - bi : r));
    return r < 0 ? q - 1u : (r >= bi ? q + 1u : q);
}

// One tick over the region's items (plan_region).
// FASTDIV: the items' index arithmetic by float reciprocals (divmod_rcp) when
// the region has fewer than 2^22 items (always, for grids the planner makes:
// items ~ max(word columns x rows, threads x items per thread)); integer
// division otherwise, and in kernels with split jobs.
template <uint32_t FLAGS, bool COH, bool GLOBAL, uint32_t SHAPE, bool FASTDIV>
__device__ __forceinline__ void tick_items(const JobGeo &g, const Words<GLOBAL> &S, uint32_t acur, uint32_t anext,
                           const RuleConsts &rc, const Region &r, StatAcc &acc)
{
    const bool fast = FASTDIV && r.n < (1u << 22);
    for (uint32_t it = threadIdx.x; it < r.n; it += blockDim.x) {
        uint32_t wr, yr, rest, zc;
        if (fast) {
            rest = divmod_rcp(it, r.nw, __builtin_amdgcn_rcpf((float)r.nw), wr);
            zc = divmod_rcp(rest, r.ny, __builtin_amdgcn_rcpf((float)r.ny), yr);
        } else {
            wr = it % r.nw;
            rest = it / r.nw;
            yr = rest % r.ny;
            zc = rest / r.ny;
        }
        const int w = (int)(r.w0 + wr);
        const int y = (int)(r.y0 + yr);
        const int z0 = (int)(r.za + zc * r.tz);
        const int z1 = min(z0 + (int)r.tz, (int)r.zb);
        if constexpr (SHAPE != kBatchShapeCube)
            run_item_erd<FLAGS, COH, GLOBAL, SHAPE == kBatchShapeErdWrap>(g, S, acur, anext, rc, w, y, z0, z1, acc);
        else run_item<FLAGS, COH, GLOBAL>(g, S, acur, anext, rc, w, y, z0, z1, acc);
    }
}

// flags: F|E|C of a cube job, F|E of an ERD job.  The shape is a template
// parameter (a kernel instantiation per shape): with both shapes' tick code in
// one kernel this compiler fails ("illegal VGPR to SGPR copy").
template <bool COH, bool GLOBAL, uint32_t SHAPE, bool FASTDIV>
__device__ __forceinline__ void run_tick(uint32_t flags, const JobGeo &g, const Words<GLOBAL> &S, uint32_t acur, uint32_t anext,
                         const RuleConsts &rc, const Region &r, StatAcc &acc)
{
    if constexpr (SHAPE != kBatchShapeCube) {
        switch (flags & 3u) {
        case 1: tick_items<1, COH, GLOBAL, SHAPE, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
        case 2: tick_items<2, COH, GLOBAL, SHAPE, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
        case 3: tick_items<3, COH, GLOBAL, SHAPE, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
        default: tick_items<0, COH, GLOBAL, SHAPE, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
        }
        return;
    }
    switch (flags & 7u) {
    case 1: tick_items<1, COH, GLOBAL, kBatchShapeCube, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
    case 2: tick_items<2, COH, GLOBAL, kBatchShapeCube, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
    case 3: tick_items<3, COH, GLOBAL, kBatchShapeCube, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
    case 4: tick_items<4, COH, GLOBAL, kBatchShapeCube, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
    case 5: tick_items<5, COH, GLOBAL, kBatchShapeCube, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
    case 6: tick_items<6, COH, GLOBAL, kBatchShapeCube, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
    case 7: tick_items<7, COH, GLOBAL, kBatchShapeCube, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
    default: tick_items<0, COH, GLOBAL, kBatchShapeCube, FASTDIV>(g, S, acur, anext, rc, r, acc); break;
    }
}

// Initial cells from the hash (searchSystem.cpp:264-303), natural bit order,
// for this part's planes [zlo, zhi).  A part owns the live-plane slabs of its
// planes (the first part also the z = -1 guard slab, the last part the z = Z
// one): it clears them in both live planes, then writes its live bits.
template <bool COH, bool GLOBAL>
__device__ void init_cells(const DevJob &j, const JobGeo &g, const Words<GLOBAL> &S, uint32_t acur)
{
    const uint32_t parts = max(1u, j.parts);
    const uint32_t s0 = j.part == 0u ? 0u : g.zlo + 1u;
    const uint32_t s1 = j.part + 1u >= parts ? g.Z + 2u : g.zhi + 1u;
    const uint32_t nclr = (s1 > s0 ? s1 - s0 : 0u) * g.aslab;
    for (uint32_t i = threadIdx.x; i < nclr; i += blockDim.x) {
        st_live<COH>(S.at(g.alive0 + s0 * g.aslab + i), 0u, true);
        st_live<COH>(S.at(g.alive0 + g.aplane + s0 * g.aslab + i), 0u, true);
    }
    if (j.part == 0u && threadIdx.x < 2u)  // the word before each live plane
        st_live<COH>(S.at(g.alive0 - 1u + threadIdx.x * g.aplane), 0u, true);
    __syncthreads();
    const uint32_t live = j.M <= 15u ? ~0u : 0u;
    const uint32_t base = g.W * g.Y * g.zlo;
    const uint32_t n = g.W * g.Y * (g.zhi - g.zlo);
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        const uint32_t idx = base + k;
        const uint32_t w = idx % g.W;
        const uint32_t rest = idx / g.W;
        const uint32_t y = rest % g.Y;
        const uint32_t z = rest / g.Y;
        uint32_t bits = 0u;
        if (y >= j.init.box_lo[1] && y < j.init.box_hi[1] && z >= j.init.box_lo[2] && z < j.init.box_hi[2]) {
            const uint32_t xa = max(32u * w, j.init.box_lo[0]);
            const uint32_t xb = min(min(32u * w + 32u, j.init.box_hi[0]), g.X);
            for (uint32_t x = xa; x < xb; x++)
                if (init_alive(j.init, x, y, z)) bits |= 1u << (x - 32u * w);
        }
        S.st4(4u * idx, make_uint4((j.M & 1u) ? bits : 0u, (j.M & 2u) ? bits : 0u, (j.M & 4u) ? bits : 0u,
                                   (j.M & 8u) ? bits : 0u));
        st_live<COH>(S.at(acur + w + (g.W + 1u) * ((y + 1u) + (g.Y + 2u) * (z + 1u))), bits & live, true);
    }
}

// Final cells of this part's planes in the reference packed layout
// (ceil(X/8) words per row).
__device__ void store_final(const JobGeo &g, const uint32_t *S, uint32_t *out)
{
    const uint32_t upr = (g.X + 7u) / 8u;
    const uint32_t n = upr * g.Y * g.zhi;
    for (uint32_t rw = upr * g.Y * g.zlo + threadIdx.x; rw < n; rw += blockDim.x) {
        const uint32_t xw = rw % upr;
        const uint32_t row = rw / upr;
        const uint32_t idx = (xw >> 2) + g.W * row;
        const uint32_t sh = 8u * (xw & 3u);
        uint32_t word = 0u;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t byte = (S[4u * idx + k] >> sh) & 0xFFu;
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) word |= ((byte >> i) & 1u) << (4u * i + k);
        }
        out[rw] = word;
    }
}

// A box as three words of two 16-bit fields, every field reduced by a
// minimum: w0 = min x | min y << 16, w1 = min z | ~max x << 16,
// w2 = ~max y | ~max z << 16, with ~m = 0xFFFE - m.  Coordinates are below
// 4096; an empty box (min 0x7fffffff, max -1) is all 0xFFFF.
__device__ __forceinline__ uint32_t pk_min_u16(uint32_t a, uint32_t b)
{
    typedef unsigned short us2 __attribute__((ext_vector_type(2)));
    return __builtin_bit_cast(uint32_t, __builtin_elementwise_min(__builtin_bit_cast(us2, a), __builtin_bit_cast(us2, b)));
}

__device__ __forceinline__ void box_pack(const int32_t bmin[3], const int32_t bmax[3], uint32_t w[3])
{
    uint32_t f[6];
    for (int k = 0; k < 3; k++) {
        f[k] = min((uint32_t)bmin[k], 0xFFFFu);
        f[3 + k] = 0xFFFEu - (uint32_t)bmax[k];
    }
    w[0] = f[0] | f[1] << 16;
    w[1] = f[2] | f[3] << 16;
    w[2] = f[4] | f[5] << 16;
}

__device__ __forceinline__ void box_unpack(const uint32_t w[3], int32_t bmin[3], int32_t bmax[3])
{
    const uint32_t f[6] = {w[0] & 0xFFFFu, w[0] >> 16, w[1] & 0xFFFFu, w[1] >> 16, w[2] & 0xFFFFu, w[2] >> 16};
    for (int k = 0; k < 3; k++) {
        bmin[k] = f[k] == 0xFFFFu ? 0x7fffffff : (int32_t)f[k];
        bmax[k] = (int32_t)0xFFFEu - (int32_t)f[3 + k];  // 0xFFFF: -1
    }
}

// The wave's box in the packed form: six minimum / maximum reductions, each
// step one fused DPP instruction (v_min_i32_dpp), then packed from the
// wave-uniform results with scalar instructions.  Reducing the three packed
// words instead takes three instructions a step (v_pk_min_u16 takes no DPP
// operand) plus the per-lane packing.
__device__ __forceinline__ void wave_box(const StatAcc &acc, uint32_t w[3])
{
    int32_t mn[3], mx[3];
    for (int k = 0; k < 3; k++) {
        mn[k] = wave_min_i32(acc.bmin[k]);
        mx[k] = wave_max_i32(acc.bmax[k]);
    }
    box_pack(mn, mx, w);
}

// The same reductions over lanes 0..15 only (one DPP row: the per-wave
// partials of a block of at most 16 waves): the quad and row rotation steps
// of wave_reduce_dpp, without the two row broadcasts; read from lane 0.
template <typename Op>
__device__ __forceinline__ int32_t row16_reduce_dpp(int32_t v, int32_t ident, Op op)
{
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x124, 0xF, 0xF, false));  // row_ror:4
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x128, 0xF, 0xF, false));  // row_ror:8
    return __builtin_amdgcn_readlane(v, 0);
}

__device__ __forceinline__ uint32_t row16_pk_min(uint32_t v)
{
    return (uint32_t)row16_reduce_dpp((int32_t)v, -1, [](int32_t x, int32_t y) {
        return (int32_t)pk_min_u16((uint32_t)x, (uint32_t)y);
    });
}

// Three sums and three packed minima over lanes 0..15, step by step across
// the six values, so that no DPP read waits on the instruction just before it.
__device__ __forceinline__ void row16_sums_pk_mins(uint32_t s[3], uint32_t m[3])
{
#define CAVE_ROW16_STEP(ctl)                                                                                          \
    for (int k = 0; k < 3; k++) {                                                                                     \
        s[k] += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)s[k], ctl, 0xF, 0xF, false);                           \
        m[k] = pk_min_u16(m[k], (uint32_t)__builtin_amdgcn_update_dpp(-1, (int)m[k], ctl, 0xF, 0xF, false));        \
    }
    CAVE_ROW16_STEP(0xB1)   // quad_perm [1,0,3,2]
    CAVE_ROW16_STEP(0x4E)   // quad_perm [2,3,0,1]
    CAVE_ROW16_STEP(0x124)  // row_ror:4
    CAVE_ROW16_STEP(0x128)  // row_ror:8
#undef CAVE_ROW16_STEP
    for (int k = 0; k < 3; k++) {
        s[k] = (uint32_t)__builtin_amdgcn_readlane((int)s[k], 0);
        m[k] = (uint32_t)__builtin_amdgcn_readlane((int)m[k], 0);
    }
}

// Block min / max of the bounding box; wave 0 (thread 0 among it) receives
// it.  scratch: 3 words per wave.
__device__ __forceinline__ void block_bbox(const StatAcc &acc, uint32_t *scratch, int32_t bmin[3], int32_t bmax[3])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    uint32_t w[3];
    wave_box(acc, w);
    if (lane == 0)
        for (int k = 0; k < 3; k++) scratch[3 * wave + k] = w[k];
    __syncthreads();
    if (wave == 0) {  // lane i < 16 reduces wave i's partials (nw <= 16)
        // every lane loads (lanes past nw read a stale slot and drop it), so
        // the three loads are issued together
        const bool has = lane < nw;
        const uint32_t *p = scratch + 3 * (lane & 15);
        uint32_t pv[3];
        for (int k = 0; k < 3; k++) pv[k] = p[k];
        for (int k = 0; k < 3; k++) w[k] = row16_pk_min(has ? pv[k] : 0xFFFFFFFFu);
        box_unpack(w, bmin, bmax);
    }
}

// Block sums of the dying, born and surviving counters and the block's box,
// with one barrier: wave 0 (thread 0 among it) receives them.  scratch: 3 + 3
// words per wave.
__device__ __forceinline__ void block_stats(const StatAcc &acc, uint32_t *ssum, uint32_t *sbox, uint32_t tot[3],
                                            int32_t bmin[3], int32_t bmax[3])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    const uint32_t a = wave_sum_u32(acc.dying), b = wave_sum_u32(acc.born), c = wave_sum_u32(acc.surv);
    uint32_t w[3];
    wave_box(acc, w);
    if (lane == 0) {
        ssum[3 * wave] = a;
        ssum[3 * wave + 1] = b;
        ssum[3 * wave + 2] = c;
        for (int k = 0; k < 3; k++) sbox[3 * wave + k] = w[k];
    }
    __syncthreads();
    if (wave == 0) {  // lane i < 16 reduces wave i's partials (nw <= 16)
        // every lane loads (lanes past nw read a stale slot and drop it), so
        // the six loads are issued together
        const bool has = lane < nw;
        const uint32_t *q = ssum + 3 * (lane & 15);
        const uint32_t *p = sbox + 3 * (lane & 15);
        for (int k = 0; k < 3; k++) {
            tot[k] = q[k];
            w[k] = p[k];
        }
        for (int k = 0; k < 3; k++) {
            tot[k] = has ? tot[k] : 0u;
            w[k] = has ? w[k] : 0xFFFFFFFFu;
        }
        row16_sums_pk_mins(tot, w);
        box_unpack(w, bmin, bmax);
    }
}

// The region of the next tick (thread 0): the union u of the live boxes
// known so far, grown by `grow` cells and clipped to the grid, or the whole
// grid when `full`; re-planned only when it changed from `cur` (the region's
// box, loaded by the caller with u).  `fresh`: r holds no region yet.
// Selects, not branches: this runs on one thread while the block waits.
__device__ __forceinline__ void set_region(Region &r, const Box &u, const Box &cur, int grow, const DevJob &j,
                                           const JobGeo &g, bool full, bool fresh)
{
    const int32_t dim[3] = {(int32_t)j.X, (int32_t)j.Y, (int32_t)j.Z};
    const bool empty = (u.hi[0] < u.lo[0]) | (u.hi[1] < u.lo[1]) | (u.hi[2] < u.lo[2]);
    Box b;
    bool same = !fresh;
    for (int k = 0; k < 3; k++) {
        const int32_t lo = max(0, u.lo[k] - grow), hi = min(dim[k] - 1, u.hi[k] + grow);
        b.lo[k] = full ? 0 : (empty ? 0 : lo);
        b.hi[k] = full ? dim[k] - 1 : (empty ? -1 : hi);
        same = same & (b.lo[k] == cur.lo[k]) & (b.hi[k] == cur.hi[k]);
    }
    if (same) return;
    r.box = b;
    plan_region(r, g);
    r.cells = 0u;
    if (!empty || full) {
        const uint32_t x0 = 32u * ((uint32_t)b.lo[0] >> 5), x1 = min(j.X, 32u * (((uint32_t)b.hi[0] >> 5) + 1u));
        r.cells = (uint64_t)(x1 - x0) * (uint32_t)(b.hi[1] - b.lo[1] + 1) * (uint32_t)(b.hi[2] - b.lo[2] + 1);
    }
}

// The parts of a split job are consecutive queue entries, claimed in order:
// every part has been claimed once the claim counter (the work-queue counter,
// or the served ring's head) reaches `need`, the index one past the job's
// last part.  Until then a part may be waiting for partners that have not
// started because every block is busy, which is not a hang: the parts' wait
// limit only runs from there on.  (Every split job has at most a quarter of
// the resident blocks as parts, and the entries in flight always form one
// contiguous range of the queue, so the unclaimed parts are claimed as the
// blocks ahead of them finish.)
template <bool SERVE>
__device__ __forceinline__ bool parts_all_claimed(const BatchArgs &a, uint64_t need)
{
    if constexpr (SERVE) return __hip_atomic_load(&a.sd->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need;
    else return (uint64_t)__hip_atomic_load(a.next_job, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= need;
}

// Rendezvous of the parts of a multi-part job at the end of a phase (the
// init, then each tick).  Every wave first waits until its device-scope
// stores (the shared boundary planes) have completed; thread 0 then adds the
// block's partials to the phase's counter slot, arrives, and waits until
// every part has arrived, then reads the job totals.  Everything another
// block reads is written and read with device-scope atomics, and each wave's
// stores complete (vmcnt 0) before the arrival is issued, so a part that has
// seen every arrival reads the others' planes of this phase.  A wait is
// bounded: once every part of the job has been claimed (see
// parts_all_claimed), a wait longer than `wait_limit` aborts the job for all
// parts.  Returns false (to every thread) when the job was aborted.
template <bool SERVE>
__device__ __forceinline__ bool part_rendezvous(const BatchArgs &a, const DevJob &j, uint64_t need, uint32_t phase,
                                const uint32_t part_sum[3], const int32_t bmin[3], const int32_t bmax[3],
                                uint32_t tot[3], int32_t tmin[3], int32_t tmax[3], uint32_t *s_flag)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        JobSync *js = a.sync + j.sync_slot;
        const uint32_t slot = phase % 3u;
        if (phase > 0) {
            for (int k = 0; k < 3; k++) {
                __hip_atomic_fetch_add(&js->sum[slot][k], part_sum[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_max(&js->nmin[slot][k], 0x7fffffff - bmin[k], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_max(&js->xmax1[slot][k], bmax[k] + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        // the slot of phase + 1 was last read in phase - 2, before every part
        // arrived for phase - 1: the first part clears it for reuse
        if (j.part == 0u && phase >= 2u) {
            const uint32_t ns = (phase + 1u) % 3u;
            for (int k = 0; k < 3; k++) {
                __hip_atomic_store(&js->sum[ns][k], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&js->nmin[ns][k], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&js->xmax1[ns][k], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        __builtin_amdgcn_s_waitcnt(0);
        __hip_atomic_fetch_add(&js->arrive, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t target = j.parts * (phase + 1u);
        uint64_t t0 = wall_clock64();
        bool started = false;
        uint32_t ok = 1u;
        while (__hip_atomic_load(&js->arrive, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (!started && !(started = parts_all_claimed<SERVE>(a, need))) t0 = wall_clock64();
            if (__hip_atomic_load(&js->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                wall_clock64() - t0 > a.wait_limit) {
                __hip_atomic_store(&js->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0u;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        if (ok && phase > 0) {
            for (int k = 0; k < 3; k++) {
                tot[k] = __hip_atomic_load(&js->sum[slot][k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                tmin[k] = 0x7fffffff -
                          __hip_atomic_load(&js->nmin[slot][k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                tmax[k] = __hip_atomic_load(&js->xmax1[slot][k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - 1;
            }
        }
        *s_flag = ok;
    }
    __syncthreads();
    return *s_flag != 0u;
}

// The per-tick rendezvous of a split job whose totals fit one 64-bit word
// (fewer than 2^24 cells, no per-tick rows, the reference's death rule): each
// part adds surviving | born << 24 | 1 << 48 to the tick's slot with one
// device-scope atomic, then polls the slot until every part has arrived; the
// word then holds the job's surviving and born totals, which are all the
// manager needs per tick.  (The general rendezvous above waits for its stat
// atomics, then arrives, polls and reads the totals back: two more round
// trips per tick.)  The slot of phase + 1 (= phase - 2) was last read before
// every part arrived for phase - 1: part 0 clears it before it arrives.
//
// The same rendezvous also passes the parts' bounding boxes on, one tick late:
// each part max-combines its box of this tick into slot phase % 4 right after
// arriving (the atomics complete before its next arrival, at the wait for its
// stores that opens the next rendezvous), and once every part has arrived it
// reads the job's box of the previous tick from slot (phase - 1) % 4 into
// pmin / pmax (empty for phase 1).  Part 0 clears slot (phase + 1) % 4, last
// read before every part arrived for phase - 1, before it arrives.
template <bool SERVE>
__device__ __forceinline__ bool part_rendezvous_tick(const BatchArgs &a, const DevJob &j, uint64_t need, uint32_t phase,
                                     uint32_t born, uint32_t surv, const int32_t bmin[3], const int32_t bmax[3],
                                     uint32_t *born_tot, uint32_t *surv_tot, int32_t pmin[3], int32_t pmax[3],
                                     uint32_t *s_flag)
{
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (threadIdx.x == 0) {
        JobSync *js = a.sync + j.sync_slot;
        if (j.part == 0u) {
            if (phase >= 2u)
                __hip_atomic_store(&js->tick[(phase + 1u) % 3u], (uint64_t)0, __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
            const uint32_t ls = (phase + 1u) % 4u;
            for (int k = 0; k < 3; k++) {
                __hip_atomic_store(&js->lag_nmin[ls][k], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(&js->lag_xmax1[ls][k], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __builtin_amdgcn_s_waitcnt(0);
        }
        uint64_t *w = &js->tick[phase % 3u];
        const uint64_t add = (uint64_t)surv | ((uint64_t)born << 24) | ((uint64_t)1 << 48);
        uint64_t v = __hip_atomic_fetch_add(w, add, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + add;
        {
            const uint32_t ls = phase % 4u;
            for (int k = 0; k < 3; k++) {
                __hip_atomic_fetch_max(&js->lag_nmin[ls][k], 0x7fffffff - bmin[k], __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_max(&js->lag_xmax1[ls][k], bmax[k] + 1, __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        uint64_t t0 = wall_clock64();
        bool started = false;
        uint32_t ok = 1u;
        while ((uint32_t)(v >> 48) < j.parts) {
            if (!started && !(started = parts_all_claimed<SERVE>(a, need))) t0 = wall_clock64();
            if (__hip_atomic_load(&js->abort, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                wall_clock64() - t0 > a.wait_limit) {
                __hip_atomic_store(&js->abort, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = 0u;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
            v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        *surv_tot = (uint32_t)(v & 0xFFFFFFu);
        *born_tot = (uint32_t)((v >> 24) & 0xFFFFFFu);
        for (int k = 0; k < 3; k++) {
            pmin[k] = 0x7fffffff;
            pmax[k] = -1;
        }
        if (ok && phase >= 2u) {
            const uint32_t ls = (phase - 1u) % 4u;
            for (int k = 0; k < 3; k++) {
                pmin[k] = 0x7fffffff -
                          __hip_atomic_load(&js->lag_nmin[ls][k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                pmax[k] = __hip_atomic_load(&js->lag_xmax1[ls][k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) - 1;
            }
        }
        *s_flag = ok;
    }
    __syncthreads();
    return *s_flag != 0u;
}

// Block-wide scratch of the batched kernels.
struct BlockShared {
    uint32_t sum[3 * 16];
    uint32_t box[3 * 16];
    Region reg;        // the cells of the next tick (thread 0 writes it between ticks)
    Box u;             // the union of the live boxes so far (thread 0)
    uint64_t computed;       // cells of the regions so far, whole job (thread 0)
    CaveGridInfo info;       // the job's manager state (thread 0)
    CaveTickStats last;      // its last tick's row (thread 0)
    uint32_t done;
    uint32_t ok;
    uint64_t claim;
    DevJob job;  // k_serve: the claimed ring entry, copied from host memory
};

// What thread 0's manager reads between ticks, loaded before the statistics
// barrier.
struct MgrState {
    Box u, cur;  // the union of the live boxes so far; the region's box
    uint64_t computed, tick_cells;
    uint32_t died, born, max_alive;
};

// Runs one queue entry -- one simulation, or one part of a split one -- with
// the whole block: initial cells, the ticks with the search manager fused in
// (searchManager.slang:98-149: totals, max / final alive, early death, tick
// budget), the result record (part 0) and the final cells.  SERVE: the result
// and completion flag go to host-coherent memory, after every store of the
// simulation is visible to the host.
// FAST_SPLIT: split jobs may take the one-word per-tick rendezvous.
// SHAPE: the jobs' cell shape (kBatchShape*; make_job keeps F|E of ERD jobs).
template <bool SERVE, bool FAST_SPLIT, uint32_t SHAPE>
__device__ __forceinline__ void run_entry(const BatchArgs &a, const DevJob &j, uint32_t *lds, uint32_t *arena, BlockShared &sh)
{
    const uint64_t t_start = wall_clock64();
    const JobGeo g = job_geo(j);
    const RuleConsts rc = rule_consts_for(j.flags, j.M, j.born_lut, j.surv_lut);
    const bool multi = j.parts > 1u;
    // the claim counter value at which every part of this job has been claimed
    const uint64_t need = sh.claim + (uint64_t)(j.parts - j.part);
    uint32_t *S = j.tier == kTierOwnRegion ? reinterpret_cast<uint32_t *>(j.region) : arena;
    {   // the same in every lane: kept in scalar registers for the global tier's accesses
        const uint64_t sa = reinterpret_cast<uint64_t>(S);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)sa);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(sa >> 32));
        S = reinterpret_cast<uint32_t *>((uint64_t)hi << 32 | lo);
    }
    uint32_t acur = g.alive0, anext = g.alive0 + g.aplane;
    const Words<false> SL{lds};
    const Words<true> SG{(gword *)S};  // generic -> global address space
    if (j.tier == kTierLds) init_cells<false>(j, g, SL, acur);
    else if (multi) init_cells<true>(j, g, SG, acur);
    else init_cells<false>(j, g, SG, acur);
    // the manager's state and the last tick's row: thread 0's, kept in LDS
    CaveGridInfo &info = sh.info;
    CaveTickStats &last = sh.last;
    if (threadIdx.x == 0) {
        info = CaveGridInfo{};
        last = CaveTickStats{};
    }
    const uint32_t cells = j.X * j.Y * j.Z;
    uint32_t t3[3] = {0u, 0u, 0u};
    int32_t b3[3] = {0, 0, 0}, c3[3] = {0, 0, 0};
    bool ok = true;
    // Active region (exact compute-skip).  A cell that is 0 and has no live
    // neighbour stays 0 and counts as dead, so a tick need only compute the
    // cells that are not 0 or have a live neighbour; the stores it leaves out
    // would have written the values already there.  Every non-zero cell lies
    // in the union U of the boxes of the live cells of the ticks so far (the
    // initial box included): a cell is non-zero only for the M - 1 ticks
    // after it was live.  A tick therefore computes U grown by one cell.  The
    // live plane a tick writes was last written two ticks earlier, inside U,
    // so no stale live bit is left outside what it rewrites.  Split jobs that
    // take the one-word rendezvous learn the job's box one tick late; as a
    // box grows by at most a cell per tick, U grown by two cells covers them.
    // ERD jobs with wrap (and CAVE_SIM_FULL_GRID) compute every cell.
    const bool full = SHAPE == kBatchShapeErdWrap || (j.sim_flags & CAVE_SIM_FULL_GRID) != 0u;
    if (threadIdx.x == 0) {  // U, kept in LDS between ticks
        const uint32_t dim[3] = {j.X, j.Y, j.Z};
        Box u;
        for (int k = 0; k < 3; k++) {
            u.lo[k] = (int32_t)j.init.box_lo[k];
            u.hi[k] = (int32_t)min(j.init.box_hi[k], dim[k]) - 1;
        }
        sh.u = u;
        set_region(sh.reg, u, u, 1, j, g, full, true);
        sh.computed = 0u;
    }
    if (multi) ok = part_rendezvous<SERVE>(a, j, need, 0u, t3, b3, c3, t3, b3, c3, &sh.ok);
    else __syncthreads();
    // Split jobs whose per-tick totals fit one word take the one-atomic
    // per-tick rendezvous; their dying totals and the last tick's bounding box
    // are combined once, after the last tick.  Only kernels built with
    // FAST_SPLIT have this path: compiled in, it costs the one-block-per-job
    // code ~6% (49^3 search), so batches without split jobs run a kernel
    // without it.
    bool packed = false;
    if constexpr (FAST_SPLIT)
        packed = multi && !a.per_tick && !(j.sim_flags & CAVE_SIM_LEGACY_DEATH) && cells < (1u << 24) &&
                 j.parts < (1u << 16);
    uint32_t part_dying_all = 0u, part_dying_last = 0u;  // thread 0, packed jobs
    uint32_t ticks_run = 0;  // the same in every thread and every part
    StatAcc acc;
    for (uint32_t t = 0; ok && t < j.ticks; t++) {
        acc = StatAcc{};
        Region reg;  // the same in every thread: scalar registers, from two 16-byte LDS reads
        {
            const uint4 p0 = *reinterpret_cast<const uint4 *>(&sh.reg.w0);
            const uint4 p1 = *reinterpret_cast<const uint4 *>(&sh.reg.za);
            reg.w0 = __builtin_amdgcn_readfirstlane(p0.x);
            reg.nw = __builtin_amdgcn_readfirstlane(p0.y);
            reg.y0 = __builtin_amdgcn_readfirstlane(p0.z);
            reg.ny = __builtin_amdgcn_readfirstlane(p0.w);
            reg.za = __builtin_amdgcn_readfirstlane(p1.x);
            reg.zb = __builtin_amdgcn_readfirstlane(p1.y);
            reg.tz = __builtin_amdgcn_readfirstlane(p1.z);
            reg.n = __builtin_amdgcn_readfirstlane(p1.w);
        }
        if (j.tier == kTierLds) run_tick<false, false, SHAPE, !FAST_SPLIT>(j.flags, g, SL, acur, anext, rc, reg, acc);
        else if (multi) run_tick<true, true, SHAPE, !FAST_SPLIT>(j.flags, g, SG, acur, anext, rc, reg, acc);
        else run_tick<false, true, SHAPE, !FAST_SPLIT>(j.flags, g, SG, acur, anext, rc, reg, acc);
        ticks_run++;
        acc.born -= acc.surv;  // run_item counted every cell set to M
        // The manager's inputs, loaded now, while the block's slower waves
        // finish, so thread 0's serial part after the statistics barrier
        // waits on no LDS read.  Nothing writes these words during the tick.
        // Every thread loads them (broadcast LDS reads): loaded in thread 0
        // only, or zeroed in the others, they cost the kernel more.
        MgrState m;
        m.u = sh.u;
        m.cur = sh.reg.box;
        m.tick_cells = sh.reg.cells;  // this tick's region, before thread 0 plans the next
        m.computed = sh.computed;
        m.died = info.totalNumberOfCellsThatDied;
        m.born = info.totalNumberOfCellsThatWereBorn;
        m.max_alive = info.maxAliveCount;
        uint32_t tot[3];
        int32_t bmin[3], bmax[3];
        block_stats(acc, sh.sum, sh.box, tot, bmin, bmax);
        if (packed) {
            int32_t pmin[3], pmax[3];
            if (threadIdx.x == 0) {
                part_dying_all += tot[0];
                part_dying_last = tot[0];
            }
            ok = part_rendezvous_tick<SERVE>(a, j, need, t + 1u, tot[1], tot[2], bmin, bmax, &tot[1], &tot[2], pmin,
                                             pmax, &sh.ok);
            if (!ok) break;
            if (threadIdx.x == 0) {  // the job's box of the previous tick
                for (int k = 0; k < 3; k++) {
                    m.u.lo[k] = min(m.u.lo[k], pmin[k]);
                    m.u.hi[k] = max(m.u.hi[k], pmax[k]);
                }
                sh.u = m.u;
                set_region(sh.reg, m.u, m.cur, 2, j, g, full, false);
            }
        } else {
            // every tick: whether this tick is the last is known only after the totals
            if (multi) ok = part_rendezvous<SERVE>(a, j, need, t + 1u, tot, bmin, bmax, tot, bmin, bmax, &sh.ok);
            if (!ok) break;
            if (threadIdx.x == 0) {  // the job's box of this tick
                for (int k = 0; k < 3; k++) {
                    m.u.lo[k] = min(m.u.lo[k], bmin[k]);
                    m.u.hi[k] = max(m.u.hi[k], bmax[k]);
                }
                sh.u = m.u;
                set_region(sh.reg, m.u, m.cur, 1, j, g, full, false);
            }
        }
        bool stop = false;  // thread 0
        if (threadIdx.x == 0) {
            // packed jobs learn the dying total after the last tick: it is
            // taken out of the died total and the last row then (the u32
            // totals wrap like the reference's, so the order does not matter)
            const uint32_t dying = packed ? 0u : tot[0], born = tot[1], surv = tot[2];
            const uint32_t dead = cells - dying - born - surv;
            // searchManager.slang:115-139
            info.numberOfSimulationTicksSurvived = ticks_run;
            info.totalNumberOfCellsThatDied = m.died + dead;
            info.totalNumberOfCellsThatWereBorn = m.born + born;
            info.finalAliveCount = surv;
            info.maxAliveCount = max(m.max_alive, surv);
            const bool died = (j.sim_flags & CAVE_SIM_LEGACY_DEATH) ? dead == cells : (surv == 0u && born == 0u);
            info.status = died ? 1u : (ticks_run >= j.ticks ? 2u : 0u);
            last.snap.numberOfDeadCells = dead;
            last.snap.numberOfDyingCells = dying;
            last.snap.numberOfCellsBorn = born;
            last.snap.numberOfSurvivingCells = surv;
            last.live = born + surv;
            last.changed = born + dying;
            stop = (died && !(j.sim_flags & CAVE_SIM_NO_EARLY_EXIT)) || t + 1u >= j.ticks;
            sh.done = stop ? 1u : 0u;
            sh.computed = m.computed + m.tick_cells;
        }
        __syncthreads();
        if (threadIdx.x == 0 && !packed && (stop || a.per_tick)) {
            for (int k = 0; k < 3; k++) {
                last.bbox_min[k] = bmin[k];
                last.bbox_max[k] = bmax[k];
            }
            if (a.per_tick && j.part == 0u) a.per_tick[(size_t)j.result_slot * a.per_tick_stride + t] = last;
        }
        const uint32_t tmp = acur;
        acur = anext;
        anext = tmp;
        if (sh.done) break;
    }
    // the next phase of the general rendezvous
    uint32_t phase = packed ? 1u : ticks_run + 1u;
    if (packed && ok && ticks_run > 0u) {
        int32_t bmin[3], bmax[3];
        block_bbox(acc, sh.box, bmin, bmax);
        uint32_t fin[3] = {part_dying_all, part_dying_last, 0u};
        ok = part_rendezvous<SERVE>(a, j, need, phase, fin, bmin, bmax, fin, bmin, bmax, &sh.ok);
        phase++;
        if (ok && threadIdx.x == 0) {
            const uint32_t born = last.snap.numberOfCellsBorn, surv = last.snap.numberOfSurvivingCells;
            info.totalNumberOfCellsThatDied -= fin[0];
            last.snap.numberOfDyingCells = fin[1];
            last.snap.numberOfDeadCells = cells - fin[1] - born - surv;
            last.changed = born + fin[1];
            for (int k = 0; k < 3; k++) {
                last.bbox_min[k] = bmin[k];
                last.bbox_max[k] = bmax[k];
            }
        }
    }
    if (j.has_final && ok) {
        uint32_t *dst = reinterpret_cast<uint32_t *>(j.final_addr);
        if (j.tier == kTierLds) store_final(g, lds, dst);
        else store_final(g, S, dst);
    }
    if constexpr (SERVE) {
        // the final cells of every part reach memory before part 0 reports
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
        if (multi && ok) ok = part_rendezvous<SERVE>(a, j, need, phase, t3, b3, c3, t3, b3, c3, &sh.ok);
        else __syncthreads();
    }
    if (threadIdx.x == 0 && j.part == 0u) {
        DevResult &r = a.results[j.result_slot];  // written field by field, from LDS
        r.info = info;
        // searchSystem.cpp:695-711
        r.classification = info.status == 2u ? 2u : (info.numberOfSimulationTicksSurvived > j.M + 1u ? 1u : 0u);
        r.ticks_run = info.numberOfSimulationTicksSurvived;
        r.last = last;
        r.t_start = t_start;
        r.t_end = wall_clock64();
        r.err = ok ? 0u : 1u;
        r.pad_ = 0u;
        r.computed = sh.computed;
        if constexpr (SERVE) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
            __hip_atomic_store(a.done + j.result_slot, j.seq, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    __syncthreads();
}

// SPLIT: the batch holds split jobs (FAST_SPLIT entry code).  SHAPE: its
// jobs' cell shape (a launch runs one shape).
template <bool SPLIT, uint32_t SHAPE>
__global__ void __launch_bounds__(1024) k_batch(BatchArgs a)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];  // 16-byte state accesses
    __shared__ BlockShared sh;

    uint32_t *arena = a.arena ? a.arena + (size_t)blockIdx.x * a.arena_words : nullptr;
    while (true) {
        if (threadIdx.x == 0) sh.claim = atomicAdd(a.next_job, 1u);
        __syncthreads();
        // the same in every thread: a scalar register, so the job's fields
        // load into scalar registers (the rule tables must: see apply_rule_consts)
        const uint32_t jid = __builtin_amdgcn_readfirstlane((uint32_t)sh.claim);
        __syncthreads();
        if (jid >= a.njobs) break;
        const DevJob j = a.jobs[jid];
        run_entry<false, SPLIT, SHAPE>(a, j, lds, arena, sh);
    }
}

// Claims the next ring entry for the block (thread 0).  Idle blocks compare
// the claim counter with the last tail seen on the device; only the block
// holding the poll lock reads the host's tail over the bus.  That block also
// closes the kernel -- for every block at once -- when every entry seen has
// been claimed and the host asked to stop, or no claim happened for
// idle_limit.  Entries the host publishes after that wait for the next launch.
__device__ bool serve_claim(const BatchArgs &a, uint64_t &jid)
{
    ServeDev *sd = a.sd;
    uint32_t naps = 0;
    while (true) {
        if (__hip_atomic_load(&sd->closing, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) return false;
        uint64_t h = __hip_atomic_load(&sd->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint64_t seen = __hip_atomic_load(&sd->tail_seen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (h < seen) {
            if (__hip_atomic_compare_exchange_strong(&sd->head, &h, h + 1u, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_store(&sd->last_claim, wall_clock64(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                jid = h;
                return true;
            }
            continue;
        }
        // The lock holder alone raises tail_seen and alone sets closing.  Once
        // closing is set, tail_seen never moves again and every entry below
        // it is claimed, so no block can claim a part of a split simulation
        // whose other parts a closing kernel would leave unclaimed (those
        // parts would wait for partners that never start, and the kernel
        // could not end to be relaunched).  Published tails fall on whole
        // requests, so "every entry below tail_seen claimed" means every part
        // of every request seen is running.
        // Idle blocks read the lock before trying it, so they do not keep one
        // cache line busy with read-modify-writes while split jobs rendezvous.
        uint32_t unlocked = 0u;
        if (__hip_atomic_load(&sd->poll_lock, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u &&
            __hip_atomic_compare_exchange_strong(&sd->poll_lock, &unlocked, 1u, __ATOMIC_ACQUIRE, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
            if (__hip_atomic_load(&sd->closing, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_store(&sd->poll_lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
                return false;
            }
            const uint64_t tail = __hip_atomic_load(&a.ctl->tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            const uint64_t stop = __hip_atomic_load(&a.ctl->stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            const uint64_t before = __hip_atomic_fetch_max(&sd->tail_seen, tail, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT);
            const uint64_t seen_now = before > tail ? before : tail;
            const uint64_t now_head = __hip_atomic_load(&sd->head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            bool close = false;
            if (now_head >= seen_now) {  // every entry seen is claimed
                // Idle time is a signed difference: the claim stamp may come
                // from a block whose clock read is a little ahead of this one.
                // A zero stamp (a fresh launch) starts the idle clock now.
                const uint64_t now = (uint64_t)wall_clock64();
                uint64_t last = __hip_atomic_load(&sd->last_claim, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (last == 0u) {
                    __hip_atomic_compare_exchange_strong(&sd->last_claim, &last, now, __ATOMIC_RELAXED,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    close = stop != 0u;
                } else {
                    close = stop != 0u || (int64_t)(now - last) > (int64_t)a.idle_limit;
                }
            }
            if (close) __hip_atomic_store(&sd->closing, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(&sd->poll_lock, 0u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            if (close) return false;
            if (now_head < seen_now) continue;
        }
        // back off: short naps while work may be arriving, then the longest
        // nap (~8K cycles), which keeps idle blocks' polling off the memory
        // system that split requests rendezvous through
        if (naps++ < 16u) __builtin_amdgcn_s_sleep(4);
        else __builtin_amdgcn_s_sleep(127);
    }
}

// Persistent server: blocks claim ring entries (host-coherent memory) as the
// host publishes them, run them with the batched kernel's code, and report
// each result with a completion flag.  Every block exits once the kernel is
// closed (see serve_claim).
__global__ void __launch_bounds__(1024) k_serve(BatchArgs a)
{
    extern __shared__ __attribute__((aligned(16))) uint32_t lds[];  // 16-byte state accesses
    __shared__ BlockShared sh;

    constexpr uint32_t kJobWords = sizeof(DevJob) / 4u;
    while (true) {
        if (threadIdx.x == 0) {
            uint64_t jid = 0;
            sh.claim = serve_claim(a, jid) ? jid : ~0ull;
        }
        __syncthreads();
        const uint64_t jid = sh.claim;
        if (jid == ~0ull) break;
        // the entry lives in host memory: one load per word, then every
        // thread takes the words from LDS into scalar registers
        const uint32_t *src = reinterpret_cast<const uint32_t *>(a.jobs + jid % a.ring_cap);
        if (threadIdx.x < kJobWords)
            reinterpret_cast<uint32_t *>(&sh.job)[threadIdx.x] =
                __hip_atomic_load(src + threadIdx.x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __syncthreads();
        DevJob j;
#pragma unroll
        for (uint32_t k = 0; k < kJobWords; k++)
            reinterpret_cast<uint32_t *>(&j)[k] =
                __builtin_amdgcn_readfirstlane(reinterpret_cast<const uint32_t *>(&sh.job)[k]);
        __syncthreads();
        run_entry<true, true, kBatchShapeCube>(a, j, lds, nullptr, sh);
    }
}

const void *serve_kernel() { return (const void *)k_serve; }

int launch_serve(const BatchArgs &a, uint32_t blocks, uint32_t threads, uint32_t lds_bytes, hipStream_t st)
{
    if (lds_bytes > 64u * 1024u)
        CAVE_HIP(hipFuncSetAttribute((const void *)k_serve, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
    hipLaunchKernelGGL(k_serve, dim3(blocks), dim3(threads), lds_bytes, st, a);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

const void *batch_kernel(bool split, uint32_t shape)
{
    switch (shape) {
    case kBatchShapeErd: return split ? (const void *)k_batch<true, kBatchShapeErd> : (const void *)k_batch<false, kBatchShapeErd>;
    case kBatchShapeErdWrap:
        return split ? (const void *)k_batch<true, kBatchShapeErdWrap> : (const void *)k_batch<false, kBatchShapeErdWrap>;
    default: return split ? (const void *)k_batch<true, kBatchShapeCube> : (const void *)k_batch<false, kBatchShapeCube>;
    }
}

int launch_batch(const BatchArgs &a, uint32_t blocks, uint32_t threads, uint32_t lds_bytes, hipStream_t st,
                 bool split_jobs, uint32_t shape)
{
    const void *k = batch_kernel(split_jobs, shape);
    if (lds_bytes > 64u * 1024u) CAVE_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
    void *args[] = {(void *)&a};
    CAVE_HIP(hipLaunchKernel(k, dim3(blocks), dim3(threads), args, lds_bytes, st));
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

}  // namespace cave
