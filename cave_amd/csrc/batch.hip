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
//   * 4 state bit planes, word w + W*(y + Y*z), updated in place;
//   * 2 live planes (ping-pong) with a zero guard band: rows y = -1 and Y and
//     planes z = -1 and Z exist and stay 0, word w + W*((y+1) + (Y+2)*(z+1)),
//     so neighbour reads never branch on the grid faces.
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
    uint32_t pplane;     // W * Y * Z: one state plane
    uint32_t aplane;     // W * (Y+2) * (Z+2): one guarded live plane
    uint32_t aslab;      // W * (Y+2): words per z plane of a live plane
    uint32_t last_mask;  // valid cells of the last word of a row
    uint32_t zchunks, tzc, nitems;
};

__device__ __forceinline__ JobGeo job_geo(const DevJob &j, uint32_t zchunks)
{
    JobGeo g;
    g.X = j.X;
    g.Y = j.Y;
    g.Z = j.Z;
    g.W = (j.X + 31u) / 32u;
    g.pplane = g.W * j.Y * j.Z;
    g.aslab = g.W * (j.Y + 2u);
    g.aplane = g.aslab * (j.Z + 2u);
    const uint32_t r = j.X - 32u * (g.W - 1u);
    g.last_mask = r >= 32u ? ~0u : ((1u << r) - 1u);
    g.tzc = (j.Z + zchunks - 1u) / zchunks;
    g.zchunks = (j.Z + g.tzc - 1u) / g.tzc;
    g.nitems = g.W * j.Y * g.zchunks;
    return g;
}

struct BCarry {
    Bits<5> Q;    // S1 of the plane before the last input plane + S0 of the last
    Bits<4> S1p;  // S1 of the last input plane
    uint32_t A;   // own live word of the last input plane
};

// One item: column (w, y), planes [z0, z1).  A_cur / A_next / P are offsets
// into S.  Loads of the x-neighbour words of a row always hit valid storage
// (the previous / next row or a guard row); words outside the row are masked.
template <uint32_t FLAGS>
__device__ __forceinline__ void run_item(const JobGeo &g, uint32_t *S, uint32_t acur, uint32_t anext,
                                         const RuleConsts &rc, int w, int y, int z0, int z1, StatAcc &acc)
{
    const uint32_t vmask = (uint32_t)w + 1u < g.W ? ~0u : g.last_mask;
    const uint32_t pm = w > 0 ? ~0u : 0u;
    const uint32_t nm = (uint32_t)w + 1u < g.W ? ~0u : 0u;
    // live word of row (y-1) at plane zi: acur + w + W*(y + (Y+2)*(zi+1))
    const uint32_t *A = S + acur + w + g.W * (uint32_t)y;
    uint32_t colmask = 0u;
    int zmin = 0x7fffffff, zmax = -1;

    auto step = [&](int zi, const BCarry &ci, BCarry &co) {
        RowSums rs[3];
        const uint32_t *a0 = A + g.aslab * (uint32_t)(zi + 1);
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const uint32_t *p = a0 + d * g.W;
            const uint32_t a = p[0];
            const uint32_t prev = p[-1] & pm;
            const uint32_t next = p[1] & nm;
            rs[d] = row_sums(a, shift_left_neighbor_nat(a, prev), shift_right_neighbor_nat(a, next));
        }
        const Bits<4> S0c = add3<4>(kind1<FLAGS>(rs[0]), kind0<FLAGS>(rs[1]), kind1<FLAGS>(rs[2]));
        const Bits<4> S1c = add3<4>(kind2<FLAGS>(rs[0]), kind1<FLAGS>(rs[1]), kind2<FLAGS>(rs[2]));
        co.Q = add3<5>(ci.S1p, S0c, Bits<1>{{0u}});
        co.S1p = S1c;
        co.A = rs[1].a;
        const int z = zi - 1;
        if (z >= z0) {
            const Bits<5> n = add3<5>(ci.Q, S1c, Bits<1>{{0u}});
            uint32_t tb, ts;
            apply_rule_rows<1>(&n, rc.born_lut, rc.surv_lut, &tb, &ts);
            uint32_t *P = S + (uint32_t)w + g.W * ((uint32_t)y + g.Y * (uint32_t)z);
            uint32_t pl[4] = {P[0], P[g.pplane], P[2u * g.pplane], P[3u * g.pplane]};
            const uint32_t NZ = nonzero_mask(pl);
            const uint32_t Al = ci.A;
            const uint32_t set_m = vmask & ((Al & ts) | (~NZ & tb));
            const uint32_t surv = set_m & Al;
            acc.surv += __popc(surv);
            acc.born += __popc(set_m ^ surv);
            acc.dying += __popc(NZ & ~surv);
            update_planes(pl, set_m, NZ, rc.mk);
            P[0] = pl[0];
            P[g.pplane] = pl[1];
            P[2u * g.pplane] = pl[2];
            P[3u * g.pplane] = pl[3];
            S[anext + (uint32_t)w + g.W * ((uint32_t)(y + 1) + (g.Y + 2u) * (uint32_t)(z + 1))] = set_m;
            colmask |= set_m;
            if (set_m) {
                zmin = min(zmin, z);
                zmax = max(zmax, z);
            }
        }
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

template <uint32_t FLAGS>
__device__ void tick_items(const JobGeo &g, uint32_t *S, uint32_t acur, uint32_t anext, const RuleConsts &rc,
                           StatAcc &acc)
{
    for (uint32_t it = threadIdx.x; it < g.nitems; it += blockDim.x) {
        const int w = (int)(it % g.W);
        const uint32_t rest = it / g.W;
        const int y = (int)(rest % g.Y);
        const uint32_t zc = rest / g.Y;
        const int z0 = (int)(zc * g.tzc);
        const int z1 = min(z0 + (int)g.tzc, (int)g.Z);
        run_item<FLAGS>(g, S, acur, anext, rc, w, y, z0, z1, acc);
    }
}

__device__ void run_tick(uint32_t flags, const JobGeo &g, uint32_t *S, uint32_t acur, uint32_t anext,
                         const RuleConsts &rc, StatAcc &acc)
{
    switch (flags & 7u) {
    case 1: tick_items<1>(g, S, acur, anext, rc, acc); break;
    case 2: tick_items<2>(g, S, acur, anext, rc, acc); break;
    case 3: tick_items<3>(g, S, acur, anext, rc, acc); break;
    case 4: tick_items<4>(g, S, acur, anext, rc, acc); break;
    case 5: tick_items<5>(g, S, acur, anext, rc, acc); break;
    case 6: tick_items<6>(g, S, acur, anext, rc, acc); break;
    case 7: tick_items<7>(g, S, acur, anext, rc, acc); break;
    default: tick_items<0>(g, S, acur, anext, rc, acc); break;
    }
}

// Initial cells from the hash (searchSystem.cpp:264-303), natural bit order;
// every live-plane word (guards included) is cleared first.
__device__ void init_cells(const DevJob &j, const JobGeo &g, uint32_t *S, uint32_t acur)
{
    for (uint32_t i = threadIdx.x; i < 2u * g.aplane; i += blockDim.x) S[4u * g.pplane + i] = 0u;
    __syncthreads();
    const uint32_t live = j.M <= 15u ? ~0u : 0u;
    for (uint32_t idx = threadIdx.x; idx < g.pplane; idx += blockDim.x) {
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
        S[idx] = (j.M & 1u) ? bits : 0u;
        S[g.pplane + idx] = (j.M & 2u) ? bits : 0u;
        S[2u * g.pplane + idx] = (j.M & 4u) ? bits : 0u;
        S[3u * g.pplane + idx] = (j.M & 8u) ? bits : 0u;
        S[acur + w + g.W * ((y + 1u) + (g.Y + 2u) * (z + 1u))] = bits & live;
    }
}

// Final cells in the reference packed layout (ceil(X/8) words per row).
__device__ void store_final(const JobGeo &g, const uint32_t *S, uint32_t *out)
{
    const uint32_t upr = (g.X + 7u) / 8u;
    const uint32_t n = upr * g.Y * g.Z;
    for (uint32_t rw = threadIdx.x; rw < n; rw += blockDim.x) {
        const uint32_t xw = rw % upr;
        const uint32_t row = rw / upr;
        const uint32_t idx = (xw >> 2) + g.W * row;
        const uint32_t sh = 8u * (xw & 3u);
        uint32_t word = 0u;
#pragma unroll
        for (uint32_t k = 0; k < 4; k++) {
            const uint32_t byte = (S[k * g.pplane + idx] >> sh) & 0xFFu;
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) word |= ((byte >> i) & 1u) << (4u * i + k);
        }
        out[rw] = word;
    }
}

// Block sums of three counters; thread 0 receives them.  scratch: 3 words per wave.
__device__ __forceinline__ void block_sum3(uint32_t a, uint32_t b, uint32_t c, uint32_t *scratch, uint32_t out[3])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    a = wave_sum_u32(a);
    b = wave_sum_u32(b);
    c = wave_sum_u32(c);
    if (lane == 0) {
        scratch[3 * wave] = a;
        scratch[3 * wave + 1] = b;
        scratch[3 * wave + 2] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = out[1] = out[2] = 0u;
        for (int i = 0; i < nw; i++) {
            out[0] += scratch[3 * i];
            out[1] += scratch[3 * i + 1];
            out[2] += scratch[3 * i + 2];
        }
    }
}

// Block min / max of the bounding box; thread 0 receives it.  scratch: 6 words per wave.
__device__ __forceinline__ void block_bbox(const StatAcc &acc, int32_t *scratch, int32_t bmin[3], int32_t bmax[3])
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    int32_t v[6];
    for (int k = 0; k < 3; k++) {
        v[k] = wave_min_i32(acc.bmin[k]);
        v[3 + k] = wave_max_i32(acc.bmax[k]);
    }
    if (lane == 0)
        for (int k = 0; k < 6; k++) scratch[6 * wave + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 0; k < 3; k++) {
            bmin[k] = 0x7fffffff;
            bmax[k] = -1;
        }
        for (int i = 0; i < nw; i++)
            for (int k = 0; k < 3; k++) {
                bmin[k] = min(bmin[k], scratch[6 * i + k]);
                bmax[k] = max(bmax[k], scratch[6 * i + 3 + k]);
            }
    }
}

__global__ void __launch_bounds__(1024) k_batch(BatchArgs a)
{
    extern __shared__ uint32_t lds[];
    __shared__ uint32_t s_sum[3 * 16];
    __shared__ int32_t s_box[6 * 16];
    __shared__ uint32_t s_job;
    __shared__ uint32_t s_done;

    uint32_t *arena = a.arena ? a.arena + (size_t)blockIdx.x * a.arena_words : nullptr;
    while (true) {
        if (threadIdx.x == 0) s_job = atomicAdd(a.next_job, 1u);
        __syncthreads();
        const uint32_t jid = s_job;
        __syncthreads();
        if (jid >= a.njobs) break;
        const DevJob j = a.jobs[jid];
        // z-chunks per column chosen on the host (one item per thread when possible)
        const JobGeo g = job_geo(j, max(1u, j.zchunks));
        const RuleConsts rc = rule_consts(j.M, j.born_lut, j.surv_lut);
        uint32_t acur = 4u * g.pplane, anext = 4u * g.pplane + g.aplane;
        if (j.tier == 0) init_cells(j, g, lds, acur);
        else init_cells(j, g, arena, acur);
        CaveGridInfo info{};
        CaveTickStats last{};
        const uint32_t cells = j.X * j.Y * j.Z;
        __syncthreads();
        for (uint32_t t = 0; t < j.ticks; t++) {
            StatAcc acc;
            if (j.tier == 0) run_tick(j.flags, g, lds, acur, anext, rc, acc);
            else run_tick(j.flags, g, arena, acur, anext, rc, acc);
            uint32_t tot[3];
            block_sum3(acc.dying, acc.born, acc.surv, s_sum, tot);
            if (threadIdx.x == 0) {
                const uint32_t dying = tot[0], born = tot[1], surv = tot[2];
                const uint32_t dead = cells - dying - born - surv;
                // searchManager.slang:115-139
                info.numberOfSimulationTicksSurvived++;
                info.totalNumberOfCellsThatDied += dead;
                info.totalNumberOfCellsThatWereBorn += born;
                info.finalAliveCount = surv;
                info.maxAliveCount = max(info.maxAliveCount, surv);
                const bool died = (j.sim_flags & CAVE_SIM_LEGACY_DEATH) ? dead == cells : (surv == 0u && born == 0u);
                if (died) info.status = 1u;
                else if (info.numberOfSimulationTicksSurvived >= j.ticks) info.status = 2u;
                else info.status = 0u;
                last.snap.numberOfDeadCells = dead;
                last.snap.numberOfDyingCells = dying;
                last.snap.numberOfCellsBorn = born;
                last.snap.numberOfSurvivingCells = surv;
                last.live = born + surv;
                last.changed = born + dying;
                const bool stop = (info.status == 1u && !(j.sim_flags & CAVE_SIM_NO_EARLY_EXIT)) || t + 1u >= j.ticks;
                s_done = stop ? 1u : 0u;
            }
            __syncthreads();
            // the bounding box is reduced only when someone reads it
            if (s_done || a.per_tick) {
                int32_t bmin[3], bmax[3];
                block_bbox(acc, s_box, bmin, bmax);
                if (threadIdx.x == 0) {
                    for (int k = 0; k < 3; k++) {
                        last.bbox_min[k] = bmin[k];
                        last.bbox_max[k] = bmax[k];
                    }
                    if (a.per_tick) a.per_tick[(size_t)j.result_slot * a.per_tick_stride + t] = last;
                }
            }
            const uint32_t tmp = acur;
            acur = anext;
            anext = tmp;
            const bool done = s_done;
            __syncthreads();
            if (done) break;
        }
        if (threadIdx.x == 0) {
            DevResult r;
            r.info = info;
            // searchSystem.cpp:695-711
            r.classification = info.status == 2u ? 2u : (info.numberOfSimulationTicksSurvived > j.M + 1u ? 1u : 0u);
            r.ticks_run = info.numberOfSimulationTicksSurvived;
            r.last = last;
            a.results[j.result_slot] = r;
        }
        if (j.final_slot) {
            uint32_t *dst = a.final_cells + a.final_offsets[j.final_slot - 1u];
            if (j.tier == 0) store_final(g, lds, dst);
            else store_final(g, arena, dst);
        }
        __syncthreads();
    }
}

int launch_batch(const BatchArgs &a, uint32_t blocks, uint32_t threads, uint32_t lds_bytes, hipStream_t st)
{
    if (lds_bytes > 64u * 1024u)
        CAVE_HIP(hipFuncSetAttribute((const void *)k_batch, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
    hipLaunchKernelGGL(k_batch, dim3(blocks), dim3(threads), lds_bytes, st, a);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

}  // namespace cave
