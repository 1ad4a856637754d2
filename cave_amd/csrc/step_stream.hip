// step_stream.hip -- bit-sliced z-streaming step kernel for one large grid.
//
// Cell layouts (4 bits per cell either way):
//   * planar (default): a row is split into 32-cell groups; a group is one
//     128-bit word holding the four bit planes of its 32 cells (word k, bit c
//     = bit k of the state of cell 32g+c).  The reference's nibble layout is
//     restored at the API boundary (upload / download / init).
//   * nibble: the reference layout itself (8 nibbles per 32-bit word), rows
//     padded to whole 32-cell groups; converted to planes in registers after
//     the load and back before the store.
//
// Work decomposition (one tick per launch):
//   * a lane owns a 32-cell x-group (one 128-bit load / store) of RPL
//     consecutive y-rows, and walks a z-range of TZ planes;
//   * LPR lanes cover one x-row.  An x-row of up to 512 cells sits in one
//     16-lane DPP row and its x-neighbour live bits cross lanes with row_shr /
//     row_shl DPP moves; rows of 513..2048 cells take 32 or 64 lanes, and the
//     bits cross with whole-wave wave_shr / wave_shl DPP moves;
//   * the (64 / LPR) row slots of a wave are consecutive y blocks; the rows
//     just above / below a slot come from the neighbouring slot with a lane
//     shuffle, and only the wave's two edge slots load a halo row (a wave of
//     one slot loads both);
//   * each wave addresses its z slab (its planes plus one halo plane per
//     side) through buffer descriptors based at the slab, so 32-bit offsets
//     reach every plane whatever the grid's size;
//   * along z the kernel keeps the per-row partial sums S0 / S1 of the last
//     planes in registers, so each group is loaded once per tick (plus the
//     y / z halo of the wave tile) and stored once.  The z loop is unrolled by
//     two with two alternating register sets, so nothing is copied between
//     iterations.
// Neighbour counts, the rule and the state update are bit-sliced (32 cells
// per 32-bit instruction).  Statistics are reduced per lane, then per wave
// with cross-lane ops, then per block: one partial per block per tick, no
// atomics.
#include "bitslice.h"
#include "step_kernels.h"

namespace cave {

struct StreamArgs {
    uint32_t X, Y, Z;
    uint32_t G;          // 32-cell groups per row
    uint32_t lpr_log2;   // lanes per x-row = 1 << lpr_log2
    uint32_t rows_per_wave;
    uint32_t tz;
    uint32_t ytiles;
    uint32_t total_waves;
    uint32_t last_mask;  // valid cells of the last group of a row (in the layout's bit order)
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

__device__ __forceinline__ uint32_t dpp_wave_from_prev_lane(uint32_t v)
{   // wave_shr:1 -- lane i receives lane i-1 of the wave (0 at lane 0)
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, true);
}

__device__ __forceinline__ uint32_t dpp_wave_from_next_lane(uint32_t v)
{   // wave_shl:1 -- lane i receives lane i+1 of the wave (0 at lane 63)
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x130, 0xF, 0xF, true);
}

// How a lane reaches the groups next to its own along x.
enum : int {
    kXRows = 0,     // several x-rows per 16-lane DPP row: row shifts, masked at the x-row ends
    kXFullRow = 1,  // one x-row per 16-lane DPP row: row shifts, whose zero fill masks the ends
    kXWave32 = 2,   // one x-row per 32 lanes: whole-wave shifts, masked at the x-row ends
    kXWave64 = 3,   // one x-row per wave: whole-wave shifts, zero-filled at lanes 0 / 63; the wave
                    // holds one row slot, whose lanes load both y halo rows
};

// Lowest / highest cell index (0..31) set in a mask.
template <bool PLANAR>
__device__ __forceinline__ void mask_extent(uint32_t m, int &lo, int &hi)
{
    if (PLANAR) {
        lo = (int)__builtin_ctz(m);
        hi = 31 - (int)__builtin_clz(m);
        return;
    }
    lo = 64;  // interleaved order: bit 4i+j = cell 8j+i
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

constexpr uint32_t kOutOfRange = 0x80000000u;  // any row offset this large is outside every buffer
// Plane offset of a load that must read nothing: a wave's slab descriptor
// covers less than 2^30 bytes (stream_plan caps the z tile), and a row offset
// (< 2^30, or kOutOfRange) plus 2^30 neither wraps nor lands inside it.
constexpr uint32_t kNoPlane = 0x40000000u;
constexpr uint32_t kBufferDword3 = 0x00020000u;  // raw buffer resource word 3 on gfx9-family parts

template <int RPL, int NR>
struct Carry {
    uint4 w[NR];           // raw groups of the plane the step receiving this set consumes
    uint32_t P[RPL][4];    // state planes of the last input plane
    uint32_t A[RPL];       // its live masks
    uint32_t NZ[RPL];      // its non-zero masks
    Bits<4> S1p[RPL];      // S1 of the last input plane
    Bits<5> Q[RPL];        // S1 of the plane before it + S0 of the last input plane
};

// XMODE: one of kXRows .. kXWave64.  NR rows are loaded per step: the RPL
// owned rows and NH halo rows.
// WRAP (the W flag, planar layout): every axis wraps (rayMarchCompute.slang:
// 80-108, one term per offset).  Rows load y mod Y, so the rows past Y of a
// wave's last tile and the halo rows are the wrapped rows (only rows < Y are
// stored and counted); planes load z mod Z through one descriptor over the
// whole grid; the first and last group of an x-row take their outer
// neighbour cell from the row's other end.
template <uint32_t FLAGS, int RPL, bool PLANAR, int XMODE, bool WRAP = false>
struct StreamLane {
    static constexpr int NH = XMODE == kXWave64 ? 2 : 1;
    static constexpr int NR = RPL + NH;
    // Buffer resources over the two cell buffers: a 128-bit access whose byte
    // offset is at or beyond the buffer size reads zeros / is dropped by the
    // hardware range check, which is how rows and planes outside the grid
    // (y or z halo beyond the faces, groups past the row end) are handled
    // without branches.
    __amdgpu_buffer_rsrc_t rin, rout;  // based at plane zbase of the two buffers
    const StreamArgs &a;
    uint32_t lane, xg, slot, nslots, lpr, vmask;
    int y_first, z0, z1, zbase;
    bool yok[NR];
    uint32_t roff[NR];  // byte offset of the row within a z plane (kOutOfRange if none)
    uint32_t plane_bytes;
    // statistics
    uint32_t surv = 0, setm = 0, dying = 0, nrows = 0, colmask = 0;  // setm: cells set to M (born + surviving)
    uint32_t yany[RPL];
    int zmin = 0x7fffffff, zmax = -1;

    __device__ __forceinline__ uint4 load_at(int r, uint32_t zoff) const
    {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(rin, (int)(roff[r] + zoff), 0, 0);
        return make_uint4(v[0], v[1], v[2], v[3]);
    }

    // the plane a load of plane z reads: z itself (-1 and Z fall outside the
    // slab and read zeros), or z mod Z with WRAP (z is in -1..Z)
    __device__ __forceinline__ int plane_of(int z) const
    {
        if constexpr (WRAP) return z < 0 ? z + (int)a.Z : (z >= (int)a.Z ? z - (int)a.Z : z);
        return z;
    }

    __device__ __forceinline__ uint4 load(int r, int z) const
    {   // z = -1 wraps to a huge offset and z = Z lands at the slab end: both read zeros
        return load_at(r, (uint32_t)(plane_of(z) - zbase) * plane_bytes);
    }

    __device__ __forceinline__ void store(int r, int z, const uint32_t P[4]) const
    {   // WRAP: a row past Y holds a wrapped copy and is never stored
        typedef unsigned int v4u __attribute__((ext_vector_type(4)));
        const v4u v = {P[0], P[1], P[2], P[3]};
        const uint32_t ro = (WRAP && !yok[r]) ? kOutOfRange : roff[r];
        __builtin_amdgcn_raw_buffer_store_b128(v, rout, (int)(ro + (uint32_t)(z - zbase) * plane_bytes), 0, 0);
    }

    __device__ __forceinline__ void planes_of(const uint4 &w, uint32_t P[4]) const
    {
        P[0] = w.x;
        P[1] = w.y;
        P[2] = w.z;
        P[3] = w.w;
        if (!PLANAR) transpose4x4(P[0], P[1], P[2], P[3]);
    }

    // One z step of a walk in direction dz (+1 or -1): consume plane zi
    // (ci.w), load plane zi+dz into co.w when `more` (the next step consumes
    // it), finalize plane zi-dz when `fin`.  The z stencil is symmetric, so a
    // descending walk computes the same planes.
    __device__ __forceinline__ void step(int zi, int dz, bool more, bool fin, Carry<RPL, NR> &ci,
                                         Carry<RPL, NR> &co)
    {
        const RuleConsts &rc = a.rc;
        RowSums rs[NR];
        uint4 wcur[NR];
#pragma unroll
        for (int r = 0; r < NR; r++) wcur[r] = ci.w[r];
        // Unconditional: past the walk's last plane the load gets an offset
        // the range check rejects (no memory access, zeros, never consumed).
        // A conditional load would be waited on at once.
        const uint32_t zoff = more ? (uint32_t)(plane_of(zi + dz) - zbase) * plane_bytes : kNoPlane;
#pragma unroll
        for (int r = 0; r < NR; r++) co.w[r] = load_at(r, zoff);
#pragma unroll
        for (int r = 0; r < NR; r++) {
            uint32_t P[4];
            planes_of(wcur[r], P);
            const uint32_t A = live_mask_v3(P, rc.nm, rc.en);
            if (r < RPL) {
                co.P[r][0] = P[0];
                co.P[r][1] = P[1];
                co.P[r][2] = P[2];
                co.P[r][3] = P[3];
                co.A[r] = A;
                co.NZ[r] = nonzero_mask(P);
            }
            uint32_t prev, next;
            if constexpr (XMODE >= kXWave32) {
                prev = dpp_wave_from_prev_lane(A);
                next = dpp_wave_from_next_lane(A);
            } else {
                prev = dpp_from_prev_lane(A);
                next = dpp_from_next_lane(A);
            }
            if constexpr (XMODE == kXRows || XMODE == kXWave32) {  // several x-rows share the shifted lanes
                prev = xg == 0u ? 0u : prev;
                next = xg == lpr - 1u ? 0u : next;
            }
            uint32_t wrap_r = 0u;
            if constexpr (WRAP) {
                // cell X-1 (bit b of group G-1) is the left neighbour of cell 0,
                // cell 0 the right neighbour of cell X-1; lanes past G hold 0
                const int base = (int)(lane - xg);
                const uint32_t firstA = (uint32_t)__shfl((int)A, base, 64);
                const uint32_t lastA = (uint32_t)__shfl((int)A, base + (int)a.G - 1, 64);
                const uint32_t b = (a.X - 1u) & 31u;
                if (xg == 0u) prev = ((lastA >> b) & 1u) << 31;
                if (xg + 1u == a.G) wrap_r = (firstA & 1u) << b;
            }
            const uint32_t Lm = PLANAR ? shift_left_neighbor_nat(A, prev) : shift_left_neighbor_il(A, prev);
            const uint32_t Rm = (PLANAR ? shift_right_neighbor_nat(A, next) : shift_right_neighbor_il(A, next)) | wrap_r;
            rs[r] = row_sums(A, Lm, Rm);
        }

        // neighbour rows across slot boundaries
        using K1 = decltype(kind1<FLAGS>(rs[0]));
        using K2 = decltype(kind2<FLAGS>(rs[0]));
        K1 k1_up, k1_dn;
        K2 k2_up, k2_dn;
        if constexpr (NH == 2) {  // this lane loaded both halo rows: rs[RPL] above, rs[RPL + 1] below
            k1_up = kind1<FLAGS>(rs[RPL]);
            k1_dn = kind1<FLAGS>(rs[RPL + 1]);
            k2_up = kind2<FLAGS>(rs[RPL]);
            k2_dn = kind2<FLAGS>(rs[RPL + 1]);
        } else {
            const K1 k1_last = kind1<FLAGS>(rs[RPL - 1]), k1_first = kind1<FLAGS>(rs[0]), k1_h = kind1<FLAGS>(rs[RPL]);
            const K2 k2_last = kind2<FLAGS>(rs[RPL - 1]), k2_first = kind2<FLAGS>(rs[0]), k2_h = kind2<FLAGS>(rs[RPL]);
            const int src_up = ((int)lane - (int)lpr) & 63;  // slot - 1, same xg
            const int src_dn = ((int)lane + (int)lpr) & 63;  // slot + 1, same xg
            if constexpr (K1::n > 0) {
#pragma unroll
                for (int k = 0; k < K1::n; k++) {
                    const uint32_t u1 = (uint32_t)__shfl((int)k1_last.b[k], src_up, 64);
                    const uint32_t d1 = (uint32_t)__shfl((int)k1_first.b[k], src_dn, 64);
                    k1_up.b[k] = slot == 0 ? k1_h.b[k] : u1;
                    k1_dn.b[k] = slot == nslots - 1 ? k1_h.b[k] : d1;
                }
            }
            if constexpr (K2::n > 0) {
#pragma unroll
                for (int k = 0; k < K2::n; k++) {
                    const uint32_t u2 = (uint32_t)__shfl((int)k2_last.b[k], src_up, 64);
                    const uint32_t d2 = (uint32_t)__shfl((int)k2_first.b[k], src_dn, 64);
                    k2_up.b[k] = slot == 0 ? k2_h.b[k] : u2;
                    k2_dn.b[k] = slot == nslots - 1 ? k2_h.b[k] : d2;
                }
            }
        }

        // plane partial sums of plane zi
        Bits<4> S1c[RPL];
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const K1 u1 = r == 0 ? k1_up : kind1<FLAGS>(rs[r - 1 >= 0 ? r - 1 : 0]);
            const K2 u2 = r == 0 ? k2_up : kind2<FLAGS>(rs[r - 1 >= 0 ? r - 1 : 0]);
            const K1 d1 = r == RPL - 1 ? k1_dn : kind1<FLAGS>(rs[r + 1 < RPL ? r + 1 : 0]);
            const K2 d2 = r == RPL - 1 ? k2_dn : kind2<FLAGS>(rs[r + 1 < RPL ? r + 1 : 0]);
            S1c[r] = add3<4>(u2, kind1<FLAGS>(rs[r]), d2);
            if constexpr (kBoxSum<FLAGS>) {
                co.Q[r] = add2<5>(ci.S1p[r], S1c[r]);
            } else {
                const Bits<4> S0c = add3<4>(u1, kind0<FLAGS>(rs[r]), d1);
                co.Q[r] = add2<5>(ci.S1p[r], S0c);
            }
            co.S1p[r] = S1c[r];
        }

        // finalize plane z = zi - dz (uniform over the wave)
        const int z = zi - dz;
        if (fin) {
            Bits<5> n[RPL];
            uint32_t tb[RPL], ts[RPL];
#pragma unroll
            for (int r = 0; r < RPL; r++) n[r] = add2<5>(ci.Q[r], S1c[r]);
            apply_rule_consts<RPL>(n, rc, tb, ts);
            uint32_t any = 0u;
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                const uint32_t valid = yok[r] ? vmask : 0u;
                nrows += yok[r] ? 1u : 0u;
                // WRAP: a row past Y holds a wrapped copy, not zeros: its cells
                // count as nothing
                const uint32_t A = ci.A[r], NZ = WRAP ? ci.NZ[r] & valid : ci.NZ[r];
                const uint32_t set_m = valid & ((A & ts[r]) | (~NZ & tb[r]));
                const uint32_t sv = set_m & A;
                surv += __popc(sv);
                setm += __popc(set_m);
                dying += __popc(and_not(NZ, sv));
                yany[r] |= set_m;
                any |= set_m;
                uint32_t P[4] = {ci.P[r][0], ci.P[r][1], ci.P[r][2], ci.P[r][3]};
                update_planes(P, set_m, NZ, rc.mk);
                if (!PLANAR) transpose4x4(P[0], P[1], P[2], P[3]);
                store(r, z, P);
            }
            colmask |= any;
            if (any) {
                zmin = min(zmin, z);
                zmax = max(zmax, z);
            }
        }
    }
};


template <uint32_t FLAGS, int RPL, bool PLANAR, int XMODE, bool WRAP = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(RPL <= 2 ? 4 : 1))) k_step_stream(const uint4 *__restrict__ in, uint4 *__restrict__ out,
                                                     StreamArgs a, StatPart *__restrict__ partials,
                                                     const uint32_t *__restrict__ done)
{
    __shared__ int32_t red[10 * 4];
    if (done && *done) return;  // the manager already ended this simulation
    // Blocks are dealt to the 8 XCDs round-robin (block b on XCD b % 8).  When
    // the grid is a multiple of 8 blocks, XCD x takes the contiguous logical
    // blocks x*nb/8 .. (x+1)*nb/8-1, so the y and z halos a wave shares with
    // its neighbours mostly sit in the same XCD's L2.
    const uint32_t nb = gridDim.x;
    const uint32_t lblock = (nb & 7u) ? blockIdx.x : (blockIdx.x & 7u) * (nb >> 3) + (blockIdx.x >> 3);
    // readfirstlane: the wave index is uniform, and saying so keeps the tile,
    // z range and plane offsets in scalar registers
    const uint32_t gwave =
        lblock * (blockDim.x >> 6) + (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    StatAcc acc;

    if (gwave < a.total_waves) {
        using Lane = StreamLane<FLAGS, RPL, PLANAR, XMODE, WRAP>;
        constexpr int NR = Lane::NR;
        const uint32_t ytile = gwave % a.ytiles;
        const uint32_t ztile = gwave / a.ytiles;
        const int z0 = (int)(ztile * a.tz);
        const int z1 = min(z0 + (int)a.tz, (int)a.Z);
        // the wave's slab: planes zbase .. zend-1 (its own planes and the halo
        // planes inside the grid), addressed through descriptors based there
        const uint32_t plane_bytes = a.G * a.Y * 16u;
        // WRAP: the wrapped halo planes lie anywhere in the grid, so the
        // descriptors cover all of it (the plan keeps it below kMaxSlabBytes)
        const int zbase = WRAP ? 0 : max(z0 - 1, 0), zend = WRAP ? (int)a.Z : min(z1 + 1, (int)a.Z);
        const size_t slab_at = (size_t)zbase * plane_bytes;
        const int slab_bytes = (int)((uint32_t)(zend - zbase) * plane_bytes);
        Lane L{__builtin_amdgcn_make_buffer_rsrc((void *)((const char *)in + slab_at), 0, slab_bytes, kBufferDword3),
               __builtin_amdgcn_make_buffer_rsrc((void *)((char *)out + slab_at), 0, slab_bytes, kBufferDword3), a};
        L.lane = threadIdx.x & 63u;
        L.lpr = 1u << a.lpr_log2;
        L.xg = L.lane & (L.lpr - 1u);
        L.slot = L.lane >> a.lpr_log2;
        L.nslots = 64u >> a.lpr_log2;
        L.y_first = (int)(ytile * a.rows_per_wave + L.slot * RPL);
        L.z0 = z0;
        L.z1 = z1;
        L.zbase = zbase;
        const bool gvalid = L.xg < a.G;
        L.vmask = !gvalid ? 0u : (L.xg + 1u < a.G ? ~0u : a.last_mask);
        const int y_halo = L.slot == 0 ? L.y_first - 1 : (L.slot == L.nslots - 1 ? L.y_first + RPL : -1);
        L.plane_bytes = plane_bytes;
#pragma unroll
        for (int r = 0; r < NR; r++) {
            // rows RPL (and RPL + 1 with two halo rows): the halo above, then below
            const int y = r < RPL ? L.y_first + r
                                  : (Lane::NH == 2 ? (r == RPL ? L.y_first - 1 : L.y_first + RPL) : y_halo);
            L.yok[r] = gvalid && y >= 0 && y < (int)a.Y;
            if constexpr (WRAP) {
                // every row loads y mod Y (y is in -1 .. Y + rows of a wave);
                // the halo row of a slot between two others is not used and
                // loads nothing
                const bool used = r < RPL || Lane::NH == 2 || L.slot == 0 || L.slot == L.nslots - 1;
                const uint32_t yw = (uint32_t)((y % (int)a.Y + (int)a.Y) % (int)a.Y);
                L.roff[r] = gvalid && used ? (yw * a.G + L.xg) * 16u : kOutOfRange;
            } else {
                L.roff[r] = L.yok[r] ? ((uint32_t)y * a.G + L.xg) * 16u : kOutOfRange;
            }
            if (r < RPL) L.yany[r] = 0u;
        }

        Carry<RPL, NR> c0, c1;
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            for (int k = 0; k < 4; k++) c0.P[r][k] = 0u;
            c0.A[r] = c0.NZ[r] = 0u;
            for (int k = 0; k < 4; k++) c0.S1p[r].b[k] = 0u;
            for (int k = 0; k < 5; k++) c0.Q[r].b[k] = 0u;
        }
        // Odd z tiles walk downwards, so the two tiles on either side of a
        // tile boundary read the boundary planes at about the same time (both
        // at their first or both at their last steps) and the second read
        // hits L2.
        const int dz = (ztile & 1u) ? -1 : 1;
        const int zs = dz > 0 ? L.z0 : L.z1 - 1;  // first owned plane of the walk
        const int n = L.z1 - L.z0;                 // owned planes
#pragma unroll
        for (int r = 0; r < NR; r++) {
            c0.w[r] = L.load(r, zs - dz);
        }

        // steps t = 0 .. n+1 on planes zs + dz*(t-1), two per iteration with
        // alternating register sets
        int t = 0;
        for (; t + 1 <= n + 1; t += 2) {
            L.step(zs + dz * (t - 1), dz, t <= n, t >= 2, c0, c1);
            L.step(zs + dz * t, dz, t + 1 <= n, t + 1 >= 2, c1, c0);
        }
        if (t <= n + 1) L.step(zs + dz * (t - 1), dz, t <= n, t >= 2, c0, c1);

        acc.surv = L.surv;
        acc.born = L.setm - L.surv;
        acc.dying = L.dying;
        // cells that stayed 0 -> 0: every valid cell not counted as dying / born / surviving
        acc.dead = (uint32_t)__popc(L.vmask) * L.nrows - L.dying - L.setm;
        if (L.colmask) {
            int lo, hi;
            mask_extent<PLANAR>(L.colmask, lo, hi);
            acc.bmin[0] = (int)(L.xg * 32u) + lo;
            acc.bmax[0] = (int)(L.xg * 32u) + hi;
            acc.bmin[2] = L.zmin;
            acc.bmax[2] = L.zmax;
#pragma unroll
            for (int r = 0; r < RPL; r++)
                if (L.yany[r]) {
                    acc.bmin[1] = min(acc.bmin[1], L.y_first + r);
                    acc.bmax[1] = max(acc.bmax[1], L.y_first + r);
                }
        }
    }
    block_reduce_stats(acc, &partials[blockIdx.x], red);
}

// ---------------------------------------------------------------------------
// planar layout: init and conversion from / to the reference nibble layout
// ---------------------------------------------------------------------------
__global__ void k_init_planar(uint4 *cells, GridGeom g, DevInit init, uint32_t M)
{
    const uint32_t G = g.pitch / 4u;
    const size_t n = (size_t)G * g.Y * g.Z;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t gx = (uint32_t)(i % G);
        const size_t row = i / G;
        const uint32_t y = (uint32_t)(row % g.Y);
        const uint32_t z = (uint32_t)(row / g.Y);
        uint32_t bits = 0u;
        if (y >= init.box_lo[1] && y < init.box_hi[1] && z >= init.box_lo[2] && z < init.box_hi[2]) {
            const uint32_t xa = max(32u * gx, init.box_lo[0]);
            const uint32_t xb = min(min(32u * gx + 32u, init.box_hi[0]), g.X);
            for (uint32_t x = xa; x < xb; x++)
                if (init_alive(init, x, y, z)) bits |= 1u << (x - 32u * gx);
        }
        const uint32_t m = M & 15u;
        cells[i] = make_uint4((m & 1u) ? bits : 0u, (m & 2u) ? bits : 0u, (m & 4u) ? bits : 0u, (m & 8u) ? bits : 0u);
    }
}

// reference layout (ceil(X/8) words per row) -> planar groups
__global__ void k_ref_to_planar(const uint32_t *__restrict__ ref, uint4 *__restrict__ planar, GridGeom g)
{
    const uint32_t G = g.pitch / 4u, upr = (g.X + 7u) / 8u;
    const size_t n = (size_t)G * g.Y * g.Z;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t gx = (uint32_t)(i % G);
        const size_t row = i / G;
        uint32_t P[4] = {0u, 0u, 0u, 0u};
        for (uint32_t j = 0; j < 4; j++) {
            const uint32_t xw = 4u * gx + j;
            if (xw >= upr) break;
            const uint32_t w = ref[row * upr + xw];
            for (uint32_t c = 0; c < 8; c++) {
                if (8u * xw + c >= g.X) break;  // row padding is never a cell
                for (uint32_t k = 0; k < 4; k++) P[k] |= ((w >> (4u * c + k)) & 1u) << (8u * j + c);
            }
        }
        planar[i] = make_uint4(P[0], P[1], P[2], P[3]);
    }
}

// planar groups -> reference layout
__global__ void k_planar_to_ref(const uint4 *__restrict__ planar, uint32_t *__restrict__ ref, GridGeom g)
{
    const uint32_t G = g.pitch / 4u, upr = (g.X + 7u) / 8u;
    const size_t n = (size_t)upr * g.Y * g.Z;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t xw = (uint32_t)(i % upr);
        const size_t row = i / upr;
        const uint4 q = planar[row * G + xw / 4u];
        const uint32_t P[4] = {q.x, q.y, q.z, q.w};
        const uint32_t sh = 8u * (xw % 4u);
        uint32_t w = 0u;
        for (uint32_t c = 0; c < 8; c++)
            for (uint32_t k = 0; k < 4; k++) w |= ((P[k] >> (sh + c)) & 1u) << (4u * c + k);
        ref[i] = w;
    }
}

static uint32_t grid_blocks(size_t n) { return (uint32_t)std::min<size_t>((n + 255) / 256, 65535u * 4u); }

int launch_init_planar(uint32_t *cells, const GridGeom &g, const DevInit &init, uint32_t M, hipStream_t st)
{
    const size_t n = (size_t)(g.pitch / 4u) * g.Y * g.Z;
    hipLaunchKernelGGL(k_init_planar, dim3(grid_blocks(n)), dim3(256), 0, st, reinterpret_cast<uint4 *>(cells), g,
                       init, M);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

int launch_ref_to_planar(const uint32_t *ref, uint32_t *planar, const GridGeom &g, hipStream_t st)
{
    const size_t n = (size_t)(g.pitch / 4u) * g.Y * g.Z;
    hipLaunchKernelGGL(k_ref_to_planar, dim3(grid_blocks(n)), dim3(256), 0, st, ref,
                       reinterpret_cast<uint4 *>(planar), g);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

int launch_planar_to_ref(const uint32_t *planar, uint32_t *ref, const GridGeom &g, hipStream_t st)
{
    const size_t n = (size_t)((g.X + 7u) / 8u) * g.Y * g.Z;
    hipLaunchKernelGGL(k_planar_to_ref, dim3(grid_blocks(n)), dim3(256), 0, st,
                       reinterpret_cast<const uint4 *>(planar), ref, g);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
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

// Largest slab (z planes a wave addresses through one descriptor) in bytes:
// below kNoPlane, so an offset of kNoPlane past any row reads nothing.
constexpr uint64_t kMaxSlabBytes = (1ull << 30) - 1u;

bool stream_supported(const GridGeom &g)
{   // rows of <= 64 groups (one wave); any number of rows and planes (each wave
    // addresses its own slab, and a plane is at most 64 MiB: 3 planes fit a slab)
    return g.X >= 1 && g.X <= 2048 && g.pitch == 4u * ((g.X + 31u) / 32u) &&
           3ull * g.pitch * 4u * g.Y <= kMaxSlabBytes;
}

StreamPlan stream_plan(const GridGeom &g, uint32_t rpl, uint32_t tz, uint32_t waves)
{
    StreamPlan p{};
    const uint32_t G = g.pitch / 4u;
    uint32_t lpr = 1;
    while (lpr < G) lpr <<= 1;
    p.lpr = lpr;
    // rows of more than 16 groups run two rows per lane only
    p.rpl = lpr > 16u ? 2u : (rpl ? rpl : (lpr >= 8 ? 2u : 1u));
    p.waves = waves ? waves : 4u;
    const uint32_t rows_per_wave = (64u / lpr) * p.rpl;
    p.ytiles = (g.Y + rows_per_wave - 1) / rows_per_wave;
    // a wave's slab (tz planes + 2 halo planes) stays below kMaxSlabBytes
    const uint64_t plane_bytes = (uint64_t)g.pitch * 4u * g.Y;
    const uint32_t tz_cap = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(kMaxSlabBytes / plane_bytes - 2, 1u << 20));
    if (tz) {
        p.tz = std::min(tz, tz_cap);
    } else {
        // aim at >= 16 waves per CU on a 256-CU part (4 per SIMD), with z tiles of >= 8 planes
        uint32_t t = 64;
        while (t > 8 && (uint64_t)p.ytiles * ((g.Z + t - 1) / t) < 4096) t >>= 1;
        p.tz = std::min(std::min(t, g.Z), tz_cap);
    }
    p.ztiles = (g.Z + p.tz - 1) / p.tz;
    const uint32_t total = p.ytiles * p.ztiles;
    p.blocks = (total + p.waves - 1) / p.waves;
    return p;
}

template <uint32_t FLAGS, bool PLANAR, bool WRAP>
static void launch_flags(const uint4 *in, uint4 *out, const StreamArgs &a, const StreamPlan &p, StatPart *parts,
                         const uint32_t *done, hipStream_t st)
{
    const dim3 grid(p.blocks), block(64 * p.waves);
    switch (p.rpl) {
    case 1: hipLaunchKernelGGL((k_step_stream<FLAGS, 1, PLANAR, kXRows, WRAP>), grid, block, 0, st, in, out, a, parts, done); break;
    case 2:
        if (p.lpr == 64u)
            hipLaunchKernelGGL((k_step_stream<FLAGS, 2, PLANAR, kXWave64, WRAP>), grid, block, 0, st, in, out, a, parts, done);
        else if (p.lpr == 32u)
            hipLaunchKernelGGL((k_step_stream<FLAGS, 2, PLANAR, kXWave32, WRAP>), grid, block, 0, st, in, out, a, parts, done);
        else if (p.lpr == 16u)
            hipLaunchKernelGGL((k_step_stream<FLAGS, 2, PLANAR, kXFullRow, WRAP>), grid, block, 0, st, in, out, a, parts, done);
        else
            hipLaunchKernelGGL((k_step_stream<FLAGS, 2, PLANAR, kXRows, WRAP>), grid, block, 0, st, in, out, a, parts, done);
        break;
    default: hipLaunchKernelGGL((k_step_stream<FLAGS, 4, PLANAR, kXRows, WRAP>), grid, block, 0, st, in, out, a, parts, done); break;
    }
}

template <bool PLANAR, bool WRAP>
static void launch_layout(uint32_t flags, const uint4 *i4, uint4 *o4, const StreamArgs &a, const StreamPlan &p,
                          StatPart *partials, const uint32_t *done, hipStream_t st)
{
    switch (flags & 7u) {
    case 1: launch_flags<1, PLANAR, WRAP>(i4, o4, a, p, partials, done, st); break;
    case 2: launch_flags<2, PLANAR, WRAP>(i4, o4, a, p, partials, done, st); break;
    case 3: launch_flags<3, PLANAR, WRAP>(i4, o4, a, p, partials, done, st); break;
    case 4: launch_flags<4, PLANAR, WRAP>(i4, o4, a, p, partials, done, st); break;
    case 5: launch_flags<5, PLANAR, WRAP>(i4, o4, a, p, partials, done, st); break;
    case 6: launch_flags<6, PLANAR, WRAP>(i4, o4, a, p, partials, done, st); break;
    case 7: launch_flags<7, PLANAR, WRAP>(i4, o4, a, p, partials, done, st); break;
    default: launch_flags<0, PLANAR, WRAP>(i4, o4, a, p, partials, done, st); break;  // no neighbourhood: counts stay 0
    }
}

bool stream_wrap_supported(const GridGeom &g)
{   // the planar layout, and one descriptor over the whole grid
    return stream_supported(g) && g.planar && (uint64_t)g.pitch * 4u * g.Y * g.Z <= kMaxSlabBytes;
}

int launch_stream(const uint32_t *in, uint32_t *out, const GridGeom &g, const DevRule &rule, const StreamPlan &p,
                  StatPart *partials, hipStream_t st, const uint32_t *done, bool wrap)
{
    if (wrap && !stream_wrap_supported(g))
        return fail(CAVE_E_UNSUPPORTED, "the streaming kernel wraps planar grids of under 1 GiB only");
    if (p.rpl != 1 && p.rpl != 2 && p.rpl != 4) return fail(CAVE_E_INVALID, "rows per lane must be 1, 2 or 4");
    if (p.lpr > 16u && p.rpl != 2) return fail(CAVE_E_INVALID, "rows of more than 512 cells run 2 rows per lane");
    if ((uint64_t)(p.tz + 2u) * g.pitch * 4u * g.Y > kMaxSlabBytes)
        return fail(CAVE_E_INVALID, "z tile too deep: a wave's slab must stay below 1 GiB");
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
    a.last_mask = rem == 0 ? ~0u : (g.planar ? ((1u << rem) - 1u) : il_valid_mask(rem));
    a.rc = rule_consts_for(rule.flags, rule.M, rule.born_lut, rule.surv_lut);
    const uint4 *i4 = reinterpret_cast<const uint4 *>(in);
    uint4 *o4 = reinterpret_cast<uint4 *>(out);
    if (wrap) launch_layout<true, true>(rule.flags, i4, o4, a, p, partials, done, st);
    else if (g.planar) launch_layout<true, false>(rule.flags, i4, o4, a, p, partials, done, st);
    else launch_layout<false, false>(rule.flags, i4, o4, a, p, partials, done, st);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

}  // namespace cave
