// step_stream.hip -- bit-sliced z-streaming step kernel for one large grid.
//
// Work decomposition (one tick per launch):
//   * a lane owns a 32-cell x-group (four packed words, one 128-bit load) of
//     RPL consecutive y-rows, and walks a z-range of TZ planes;
//   * LPR lanes cover one x-row, so an x-row of up to 512 cells sits in one
//     16-lane DPP row: the x-neighbour live bits cross lanes with row_shr /
//     row_shl DPP moves;
//   * the (64 / LPR) row slots of a wave are consecutive y blocks; the rows
//     just above / below a slot come from the neighbouring slot with a lane
//     shuffle, and only the wave's two edge slots load a halo row;
//   * along z the kernel keeps the per-row partial sums S0 / S1 of the last
//     planes in registers, so each packed word is loaded once per tick (plus
//     the y / z halo of the wave tile) and stored once.
// Cells are converted to four bit planes right after the load; neighbour
// counts, the rule and the state update are bit-sliced (32 cells per 32-bit
// instruction), and the result is converted back to packed nibbles for the
// store.  Statistics are reduced per lane, then per wave with cross-lane ops,
// then per block; one partial per block per tick, no atomics.
#include "bitslice.h"
#include "step_kernels.h"

namespace cave {

struct StreamArgs {
    uint32_t X, Y, Z;
    uint32_t G;          // 32-cell groups per row (pitch / 4)
    uint32_t lpr_log2;   // lanes per x-row = 1 << lpr_log2
    uint32_t rows_per_wave;
    uint32_t tz;
    uint32_t ytiles;
    uint32_t total_waves;
    uint32_t last_mask;  // valid cells of the last group (interleaved order)
    RuleConsts rc;
};

__device__ __forceinline__ uint32_t dpp_from_prev_lane(uint32_t v)
{   // row_shr:1 -- lane i receives lane i-1 of its 16-lane row (0 at the row start)
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xF, 0xF, true);
}

__device__ __forceinline__ uint32_t dpp_from_next_lane(uint32_t v)
{   // row_shl:1 -- lane i receives lane i+1 of its 16-lane row (0 at the row end)
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x101, 0xF, 0xF, true);
}

// Lowest / highest cell index (0..31) set in an interleaved-order mask.
__device__ __forceinline__ void il_extent(uint32_t m, int &lo, int &hi)
{
    lo = 64;
    hi = -1;
#pragma unroll
    for (int j = 0; j < 4; j++) {
        const uint32_t mj = m & (0x11111111u << j);
        if (mj) {
            lo = min(lo, 8 * j + ((int)__builtin_ctz(mj) >> 2));
            hi = max(hi, 8 * j + ((31 - (int)__builtin_clz(mj)) >> 2));
        }
    }
}

template <uint32_t FLAGS, int RPL>
__global__ void __launch_bounds__(256) k_step_stream(const uint4 *__restrict__ in, uint4 *__restrict__ out,
                                                     StreamArgs a, StatPart *__restrict__ partials,
                                                     const uint32_t *__restrict__ done)
{
    __shared__ int32_t red[10 * 4];
    if (done && *done) return;  // the manager already ended this simulation
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t gwave = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const uint32_t lpr = 1u << a.lpr_log2;
    const uint32_t xg = lane & (lpr - 1u);
    const uint32_t slot = lane >> a.lpr_log2;
    const uint32_t nslots = 64u >> a.lpr_log2;
    StatAcc acc;

    if (gwave < a.total_waves) {
        const uint32_t ytile = gwave % a.ytiles;
        const uint32_t ztile = gwave / a.ytiles;
        const int y_first = (int)(ytile * a.rows_per_wave + slot * RPL);
        const int z0 = (int)(ztile * a.tz);
        const int z1 = min(z0 + (int)a.tz, (int)a.Z);
        const bool gvalid = xg < a.G;
        const uint32_t vmask = !gvalid ? 0u : (xg + 1u < a.G ? ~0u : a.last_mask);
        const int y_halo = slot == 0 ? y_first - 1 : (slot == nslots - 1 ? y_first + RPL : -1);
        const RuleConsts rc = a.rc;
        const size_t row_stride = a.G;           // uint4 per row
        const size_t plane_stride = (size_t)a.G * a.Y;

        auto row_ok = [&](int y, int z) {
            return gvalid && y >= 0 && y < (int)a.Y && z >= 0 && z < (int)a.Z;
        };
        auto load = [&](int y, int z) -> uint4 {
            uint4 w = make_uint4(0u, 0u, 0u, 0u);
            if (row_ok(y, z)) w = in[(size_t)z * plane_stride + (size_t)y * row_stride + xg];
            return w;
        };

        // registers carried along z, per own row
        uint32_t Pprev[RPL][4];
        uint32_t Aprev[RPL], NZprev[RPL];
        Bits<4> S1pp[RPL], S1p[RPL], S0p[RPL];
        uint32_t yany[RPL];
        uint32_t colmask = 0u;
        int zmin = 0x7fffffff, zmax = -1;
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            for (int k = 0; k < 4; k++) Pprev[r][k] = 0u;
            Aprev[r] = NZprev[r] = 0u;
            for (int k = 0; k < 4; k++) S1pp[r].b[k] = S1p[r].b[k] = S0p[r].b[k] = 0u;
            yany[r] = 0u;
        }

        // software pipeline: the words of plane zi are loaded one step ahead
        uint4 wnext[RPL + 1];
#pragma unroll
        for (int r = 0; r < RPL; r++) wnext[r] = load(y_first + r, z0 - 1);
        wnext[RPL] = load(y_halo, z0 - 1);

        for (int zi = z0 - 1; zi <= z1; zi++) {
            uint4 wcur[RPL + 1];
#pragma unroll
            for (int r = 0; r <= RPL; r++) wcur[r] = wnext[r];
            if (zi < z1) {
#pragma unroll
                for (int r = 0; r < RPL; r++) wnext[r] = load(y_first + r, zi + 1);
                wnext[RPL] = load(y_halo, zi + 1);
            }

            // planes, live masks and x-sums of the own rows and the halo row
            uint32_t Pcur[RPL][4];
            uint32_t Acur[RPL], NZcur[RPL];
            RowSums rs[RPL + 1];
#pragma unroll
            for (int r = 0; r <= RPL; r++) {
                uint32_t p0 = wcur[r].x, p1 = wcur[r].y, p2 = wcur[r].z, p3 = wcur[r].w;
                transpose4x4(p0, p1, p2, p3);
                const uint32_t P[4] = {p0, p1, p2, p3};
                const uint32_t A = live_mask(P, rc.nm, rc.en);
                if (r < RPL) {
                    Pcur[r][0] = p0; Pcur[r][1] = p1; Pcur[r][2] = p2; Pcur[r][3] = p3;
                    Acur[r] = A;
                    NZcur[r] = nonzero_mask(P);
                }
                uint32_t prev = dpp_from_prev_lane(A);
                uint32_t next = dpp_from_next_lane(A);
                prev = xg == 0u ? 0u : prev;
                next = xg == lpr - 1u ? 0u : next;
                rs[r] = row_sums(A, shift_left_neighbor_il(A, prev), shift_right_neighbor_il(A, next));
            }

            // neighbour rows across slot boundaries
            Bits<2> k1_up, k2_up, k1_dn, k2_dn;
            {
                const Bits<2> k1_last = kind1<FLAGS>(rs[RPL - 1]), k2_last = kind2<FLAGS>(rs[RPL - 1]);
                const Bits<2> k1_first = kind1<FLAGS>(rs[0]), k2_first = kind2<FLAGS>(rs[0]);
                const Bits<2> k1_h = kind1<FLAGS>(rs[RPL]), k2_h = kind2<FLAGS>(rs[RPL]);
                const int src_up = (int)lane - (int)lpr;   // slot - 1, same xg
                const int src_dn = (int)lane + (int)lpr;   // slot + 1, same xg
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    const uint32_t u1 = (uint32_t)__shfl((int)k1_last.b[k], src_up & 63, 64);
                    const uint32_t u2 = (uint32_t)__shfl((int)k2_last.b[k], src_up & 63, 64);
                    const uint32_t d1 = (uint32_t)__shfl((int)k1_first.b[k], src_dn & 63, 64);
                    const uint32_t d2 = (uint32_t)__shfl((int)k2_first.b[k], src_dn & 63, 64);
                    k1_up.b[k] = slot == 0 ? k1_h.b[k] : u1;
                    k2_up.b[k] = slot == 0 ? k2_h.b[k] : u2;
                    k1_dn.b[k] = slot == nslots - 1 ? k1_h.b[k] : d1;
                    k2_dn.b[k] = slot == nslots - 1 ? k2_h.b[k] : d2;
                }
            }

            // plane partial sums of plane zi
            Bits<4> S0c[RPL], S1c[RPL];
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                const Bits<2> u1 = r == 0 ? k1_up : kind1<FLAGS>(rs[r - 1 >= 0 ? r - 1 : 0]);
                const Bits<2> u2 = r == 0 ? k2_up : kind2<FLAGS>(rs[r - 1 >= 0 ? r - 1 : 0]);
                const Bits<2> d1 = r == RPL - 1 ? k1_dn : kind1<FLAGS>(rs[r + 1 < RPL ? r + 1 : 0]);
                const Bits<2> d2 = r == RPL - 1 ? k2_dn : kind2<FLAGS>(rs[r + 1 < RPL ? r + 1 : 0]);
                S0c[r] = add3<4>(u1, kind0<FLAGS>(rs[r]), d1);
                S1c[r] = add3<4>(u2, kind1<FLAGS>(rs[r]), d2);
            }

            // finalize plane z = zi - 1
            const int z = zi - 1;
            if (z >= z0) {
                uint32_t any = 0u;
#pragma unroll
                for (int r = 0; r < RPL; r++) {
                    const int y = y_first + r;
                    const Bits<5> n = add3<5>(S1pp[r], S0p[r], S1c[r]);
                    uint32_t tb, ts;
                    apply_rule(n, rc.born_lut, rc.surv_lut, tb, ts);
                    const bool rv = row_ok(y, z);
                    const uint32_t valid = rv ? vmask : 0u;
                    const uint32_t A = Aprev[r], NZ = NZprev[r];
                    const uint32_t zero = ~NZ;
                    const uint32_t set_m = valid & ((A & ts) | (zero & tb));
                    const uint32_t surv = set_m & A;
                    const uint32_t born = set_m & ~A;
                    const uint32_t dying = NZ & ~surv & valid;
                    acc.surv += __popc(surv);
                    acc.born += __popc(born);
                    acc.dying += __popc(dying);
                    acc.dead += __popc(valid) - __popc(dying) - __popc(surv) - __popc(born);
                    colmask |= set_m;
                    yany[r] |= set_m;
                    any |= set_m;
                    uint32_t P[4] = {Pprev[r][0], Pprev[r][1], Pprev[r][2], Pprev[r][3]};
                    update_planes(P, set_m, NZ & valid, rc.mk);
                    transpose4x4(P[0], P[1], P[2], P[3]);
                    if (rv) out[(size_t)z * plane_stride + (size_t)y * row_stride + xg] = make_uint4(P[0], P[1], P[2], P[3]);
                }
                if (any) {
                    zmin = min(zmin, z);
                    zmax = max(zmax, z);
                }
            }

            // rotate the z pipeline
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                S1pp[r] = S1p[r];
                S1p[r] = S1c[r];
                S0p[r] = S0c[r];
                for (int k = 0; k < 4; k++) Pprev[r][k] = Pcur[r][k];
                Aprev[r] = Acur[r];
                NZprev[r] = NZcur[r];
            }
        }

        // bounding box of this lane's live cells
        if (colmask) {
            int lo, hi;
            il_extent(colmask, lo, hi);
            acc.bmin[0] = (int)(xg * 32u) + lo;
            acc.bmax[0] = (int)(xg * 32u) + hi;
            acc.bmin[2] = zmin;
            acc.bmax[2] = zmax;
#pragma unroll
            for (int r = 0; r < RPL; r++)
                if (yany[r]) {
                    acc.bmin[1] = min(acc.bmin[1], y_first + r);
                    acc.bmax[1] = max(acc.bmax[1], y_first + r);
                }
        }
    }
    block_reduce_stats(acc, &partials[blockIdx.x], red);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static uint32_t il_valid_mask(uint32_t ncells)
{   // interleaved-order mask of cells 0..ncells-1 of a group
    uint32_t m = 0;
    for (uint32_t c = 0; c < ncells && c < 32; c++) m |= 1u << (4u * (c % 8u) + c / 8u);
    return m;
}

bool stream_supported(const GridGeom &g) { return g.X >= 1 && g.X <= 512 && g.pitch == 4u * ((g.X + 31u) / 32u); }

StreamPlan stream_plan(const GridGeom &g, uint32_t rpl, uint32_t tz, uint32_t waves)
{
    StreamPlan p{};
    const uint32_t G = g.pitch / 4u;
    uint32_t lpr = 1;
    while (lpr < G) lpr <<= 1;
    p.lpr = lpr;
    p.rpl = rpl ? rpl : (lpr >= 8 ? 2u : 1u);
    p.waves = waves ? waves : 4u;
    const uint32_t rows_per_wave = (64u / lpr) * p.rpl;
    p.ytiles = (g.Y + rows_per_wave - 1) / rows_per_wave;
    if (tz) {
        p.tz = tz;
    } else {
        // aim at >= 8 waves per CU on a 256-CU part, with z tiles of >= 8 planes
        uint32_t t = 64;
        while (t > 8 && (uint64_t)p.ytiles * ((g.Z + t - 1) / t) < 2048) t >>= 1;
        p.tz = std::min(t, g.Z);
    }
    p.ztiles = (g.Z + p.tz - 1) / p.tz;
    const uint32_t total = p.ytiles * p.ztiles;
    p.blocks = (total + p.waves - 1) / p.waves;
    return p;
}

template <uint32_t FLAGS>
static void launch_flags(const uint4 *in, uint4 *out, const StreamArgs &a, const StreamPlan &p, StatPart *parts,
                         const uint32_t *done, hipStream_t st)
{
    const dim3 grid(p.blocks), block(64 * p.waves);
    switch (p.rpl) {
    case 1: hipLaunchKernelGGL((k_step_stream<FLAGS, 1>), grid, block, 0, st, in, out, a, parts, done); break;
    case 2: hipLaunchKernelGGL((k_step_stream<FLAGS, 2>), grid, block, 0, st, in, out, a, parts, done); break;
    default: hipLaunchKernelGGL((k_step_stream<FLAGS, 4>), grid, block, 0, st, in, out, a, parts, done); break;
    }
}

int launch_stream(const uint32_t *in, uint32_t *out, const GridGeom &g, const DevRule &rule, const StreamPlan &p,
                  StatPart *partials, hipStream_t st, const uint32_t *done)
{
    if (p.rpl != 1 && p.rpl != 2 && p.rpl != 4) return fail(CAVE_E_INVALID, "rows per lane must be 1, 2 or 4");
    if (p.waves < 1 || p.waves > 4) return fail(CAVE_E_INVALID, "waves per block must be 1..4");
    StreamArgs a{};
    a.X = g.X;
    a.Y = g.Y;
    a.Z = g.Z;
    a.G = g.pitch / 4u;
    uint32_t l2 = 0;
    while ((1u << l2) < p.lpr) l2++;
    a.lpr_log2 = l2;
    a.rows_per_wave = (64u / p.lpr) * p.rpl;
    a.tz = p.tz;
    a.ytiles = p.ytiles;
    a.total_waves = p.ytiles * p.ztiles;
    const uint32_t rem = g.X % 32u;
    a.last_mask = rem ? il_valid_mask(rem) : ~0u;
    a.rc = rule_consts(rule.M, rule.born_lut, rule.surv_lut);
    const uint4 *i4 = reinterpret_cast<const uint4 *>(in);
    uint4 *o4 = reinterpret_cast<uint4 *>(out);
    switch (rule.flags & 7u) {
    case 1: launch_flags<1>(i4, o4, a, p, partials, done, st); break;
    case 2: launch_flags<2>(i4, o4, a, p, partials, done, st); break;
    case 3: launch_flags<3>(i4, o4, a, p, partials, done, st); break;
    case 4: launch_flags<4>(i4, o4, a, p, partials, done, st); break;
    case 5: launch_flags<5>(i4, o4, a, p, partials, done, st); break;
    case 6: launch_flags<6>(i4, o4, a, p, partials, done, st); break;
    case 7: launch_flags<7>(i4, o4, a, p, partials, done, st); break;
    default: launch_flags<0>(i4, o4, a, p, partials, done, st); break;  // no neighbourhood: counts stay 0
    }
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

}  // namespace cave
