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
// Cell storage per job (word index w + W*(y + Y*z), W = ceil(X/32)):
//   4 state bit planes (updated in place) + 2 live planes (ping-pong).
// A job whose storage fits the block's LDS keeps its whole grid in LDS for
// all of its ticks ("tier 0"); a larger job streams through a per-block
// region of a global arena ("tier 1", L2 / MALL resident for the grid sizes
// of a search).  Both tiers run the same bit-sliced tick code.
//
// Per tick, a thread owns "items": an x-word column (w, y) over a z-chunk.  It
// walks z with the plane partial sums S0 / S1 in registers (the same
// separable counting as the streaming step kernel), reading the live words of
// the 3 x 3 neighbour rows and writing the new state of its own rows.
#include "batch.h"

namespace cave {

struct JobGeo {
    uint32_t X, Y, Z, W;
    uint32_t plane;      // W * Y * Z
    uint32_t last_mask;  // valid cells of the last word of a row (natural order)
    uint32_t zchunks, tzc, nitems;
};

__device__ __forceinline__ JobGeo job_geo(const DevJob &j, uint32_t zchunks)
{
    JobGeo g;
    g.X = j.X;
    g.Y = j.Y;
    g.Z = j.Z;
    g.W = (j.X + 31u) / 32u;
    g.plane = g.W * j.Y * j.Z;
    const uint32_t r = j.X - 32u * (g.W - 1u);
    g.last_mask = r >= 32u ? ~0u : ((1u << r) - 1u);
    g.tzc = (j.Z + zchunks - 1u) / zchunks;
    g.zchunks = (j.Z + g.tzc - 1u) / g.tzc;
    g.nitems = g.W * j.Y * g.zchunks;
    return g;
}

// Live plane words of the rows around (w, y) at plane z -> RowSums.
__device__ __forceinline__ RowSums load_row(const uint32_t *__restrict__ A, const JobGeo &g, int w, int y, int z)
{
    RowSums rs;
    if (y < 0 || y >= (int)g.Y || z < 0 || z >= (int)g.Z) {
        rs.a = 0u;
        rs.lr.b[0] = rs.lr.b[1] = rs.xs.b[0] = rs.xs.b[1] = 0u;
        return rs;
    }
    const uint32_t base = (uint32_t)w + g.W * ((uint32_t)y + g.Y * (uint32_t)z);
    const uint32_t a = A[base];
    const uint32_t p = w > 0 ? A[base - 1] : 0u;
    const uint32_t n = (uint32_t)w + 1u < g.W ? A[base + 1] : 0u;
    return row_sums(a, shift_left_neighbor_nat(a, p), shift_right_neighbor_nat(a, n));
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
        const uint32_t vmask = (uint32_t)w + 1u < g.W ? ~0u : g.last_mask;
        const uint32_t *A = S + acur;
        Bits<4> S1pp{{0u, 0u, 0u, 0u}}, S1p{{0u, 0u, 0u, 0u}}, S0p{{0u, 0u, 0u, 0u}};
        uint32_t Aprev = 0u, colmask = 0u;
        int zmin = 0x7fffffff, zmax = -1;
        for (int zi = z0 - 1; zi <= z1; zi++) {
            const RowSums r0 = load_row(A, g, w, y - 1, zi);
            const RowSums r1 = load_row(A, g, w, y, zi);
            const RowSums r2 = load_row(A, g, w, y + 1, zi);
            const Bits<4> S0c = add3<4>(kind1<FLAGS>(r0), kind0<FLAGS>(r1), kind1<FLAGS>(r2));
            const Bits<4> S1c = add3<4>(kind2<FLAGS>(r0), kind1<FLAGS>(r1), kind2<FLAGS>(r2));
            const int z = zi - 1;
            if (z >= z0) {
                const Bits<5> n = add3<5>(S1pp, S0p, S1c);
                uint32_t tb, ts;
                apply_rule(n, rc.born_lut, rc.surv_lut, tb, ts);
                const uint32_t idx = (uint32_t)w + g.W * ((uint32_t)y + g.Y * (uint32_t)z);
                uint32_t P[4] = {S[idx], S[g.plane + idx], S[2u * g.plane + idx], S[3u * g.plane + idx]};
                const uint32_t NZ = nonzero_mask(P);
                const uint32_t Al = Aprev;
                const uint32_t set_m = vmask & ((Al & ts) | (~NZ & tb));
                const uint32_t surv = set_m & Al;
                const uint32_t born = set_m & ~Al;
                const uint32_t dying = NZ & ~surv;
                acc.surv += __popc(surv);
                acc.born += __popc(born);
                acc.dying += __popc(dying);
                acc.dead += __popc(vmask) - __popc(dying) - __popc(set_m);
                update_planes(P, set_m, NZ, rc.mk);
                S[idx] = P[0];
                S[g.plane + idx] = P[1];
                S[2u * g.plane + idx] = P[2];
                S[3u * g.plane + idx] = P[3];
                S[anext + idx] = set_m;
                if (set_m) {
                    colmask |= set_m;
                    zmin = min(zmin, z);
                    zmax = max(zmax, z);
                }
            }
            S1pp = S1p;
            S1p = S1c;
            S0p = S0c;
            Aprev = r1.a;
        }
        if (colmask) {
            acc.bmin[0] = min(acc.bmin[0], 32 * w + (int)__builtin_ctz(colmask));
            acc.bmax[0] = max(acc.bmax[0], 32 * w + 31 - (int)__builtin_clz(colmask));
            acc.bmin[1] = min(acc.bmin[1], y);
            acc.bmax[1] = max(acc.bmax[1], y);
            acc.bmin[2] = min(acc.bmin[2], zmin);
            acc.bmax[2] = max(acc.bmax[2], zmax);
        }
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

// Initial cells from the hash (searchSystem.cpp:264-303), natural bit order.
__device__ void init_cells(const DevJob &j, const JobGeo &g, uint32_t *S, uint32_t acur)
{
    const uint32_t live = j.M <= 15u ? ~0u : 0u;
    for (uint32_t idx = threadIdx.x; idx < g.plane; idx += blockDim.x) {
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
        S[g.plane + idx] = (j.M & 2u) ? bits : 0u;
        S[2u * g.plane + idx] = (j.M & 4u) ? bits : 0u;
        S[3u * g.plane + idx] = (j.M & 8u) ? bits : 0u;
        S[acur + idx] = bits & live;
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
            const uint32_t byte = (S[k * g.plane + idx] >> sh) & 0xFFu;
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) word |= ((byte >> i) & 1u) << (4u * i + k);
        }
        out[rw] = word;
    }
}

__global__ void __launch_bounds__(1024) k_batch(BatchArgs a)
{
    extern __shared__ uint32_t lds[];
    __shared__ int32_t red[10 * 16];
    __shared__ uint32_t s_job;
    __shared__ StatPart s_tot;
    __shared__ uint32_t s_done;

    uint32_t *arena = a.arena ? a.arena + (size_t)blockIdx.x * a.arena_words : nullptr;
    while (true) {
        if (threadIdx.x == 0) s_job = atomicAdd(a.next_job, 1u);
        __syncthreads();
        const uint32_t jid = s_job;
        __syncthreads();
        if (jid >= a.njobs) break;
        const DevJob j = a.jobs[jid];
        // items per tick: z-chunks sized so one round covers about one item per thread
        const uint32_t cols = ((j.X + 31u) / 32u) * j.Y;
        uint32_t zch = max(1u, min(j.Z, blockDim.x / max(cols, 1u)));
        const JobGeo g = job_geo(j, zch);
        uint32_t *S = j.tier == 0 ? lds : arena;
        const RuleConsts rc = rule_consts(j.M, j.born_lut, j.surv_lut);
        uint32_t acur = 4u * g.plane, anext = 5u * g.plane;
        init_cells(j, g, S, acur);
        CaveGridInfo info{};
        CaveTickStats last{};
        __syncthreads();
        for (uint32_t t = 0; t < j.ticks; t++) {
            StatAcc acc;
            run_tick(j.flags, g, S, acur, anext, rc, acc);
            block_reduce_stats(acc, &s_tot, red);
            if (threadIdx.x == 0) {
                const StatPart p = s_tot;
                // searchManager.slang:115-139
                info.numberOfSimulationTicksSurvived++;
                info.totalNumberOfCellsThatDied += p.dead;
                info.totalNumberOfCellsThatWereBorn += p.born;
                info.finalAliveCount = p.surv;
                info.maxAliveCount = max(info.maxAliveCount, p.surv);
                const bool died = (j.sim_flags & CAVE_SIM_LEGACY_DEATH)
                                      ? (uint64_t)p.dead == (uint64_t)j.X * j.Y * j.Z
                                      : (p.surv == 0u && p.born == 0u);
                if (died) info.status = 1u;
                else if (info.numberOfSimulationTicksSurvived >= j.ticks) info.status = 2u;
                else info.status = 0u;
                last.snap.numberOfDeadCells = p.dead;
                last.snap.numberOfDyingCells = p.dying;
                last.snap.numberOfCellsBorn = p.born;
                last.snap.numberOfSurvivingCells = p.surv;
                last.live = p.born + p.surv;
                last.changed = p.born + p.dying;
                for (int k = 0; k < 3; k++) {
                    last.bbox_min[k] = p.bmin[k];
                    last.bbox_max[k] = p.bmax[k];
                }
                if (a.per_tick) a.per_tick[(size_t)j.result_slot * a.per_tick_stride + t] = last;
                const bool stop = (info.status == 1u && !(j.sim_flags & CAVE_SIM_NO_EARLY_EXIT)) ||
                                  t + 1u >= j.ticks;
                s_done = stop ? 1u : 0u;
            }
            __syncthreads();
            const uint32_t tmp = acur;
            acur = anext;
            anext = tmp;
            if (s_done) break;
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
        if (j.final_slot) store_final(g, S, a.final_cells + a.final_offsets[j.final_slot - 1u]);
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
