// step_gather.hip -- cell init, the reference-form step kernel and the
// per-tick statistics reduction of the single-grid step path.
//
// The gather kernel keeps the reference's work decomposition (one lane per
// packed word = 8 cells, searchSimulationCube.slang:117-216): it is the
// simplest correct form and the kernel every faster one is checked against on
// the device.  Each lane loads the 27 words around its own once and builds an
// 8-bit live mask per word, so the per-cell neighbour test is a shift, not a
// global load.
#include "step_kernels.h"

namespace cave {

// ---------------------------------------------------------------------------
// init: writes the spawn box into a row-pitched nibble layout
// (searchSystem.cpp:264-303; pitch >= ceil(X/8) words, padding words zero).
// ---------------------------------------------------------------------------
__global__ void k_init_cells(uint32_t *cells, GridGeom g, DevInit init, uint32_t live_nibble)
{
    const size_t n = (size_t)g.pitch * g.Y * g.Z;
    for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < n;
         w += (size_t)gridDim.x * blockDim.x) {
        const uint32_t xw = (uint32_t)(w % g.pitch);
        const size_t row = w / g.pitch;
        const uint32_t y = (uint32_t)(row % g.Y);
        const uint32_t z = (uint32_t)(row / g.Y);
        uint32_t out = 0;
#pragma unroll
        for (uint32_t i = 0; i < 8; i++) {
            const uint32_t x = xw * 8u + i;
            if (x < g.X && init_alive(init, x, y, z)) out |= live_nibble << (4u * i);
        }
        cells[w] = out;
    }
}

// ---------------------------------------------------------------------------
// gather step: one lane per packed word
// ---------------------------------------------------------------------------
__device__ inline uint32_t live_mask8(uint32_t w, uint32_t M)
{
    uint32_t m = 0;
#pragma unroll
    for (uint32_t i = 0; i < 8; i++) m |= (((w >> (4u * i)) & 15u) == M ? 1u : 0u) << i;
    return m;
}

// Live bit of cell x of a row (x taken modulo X).
__device__ inline uint32_t live_bit_wrapped(const uint32_t *row, int x, uint32_t X, uint32_t M)
{
    const uint32_t xw = (uint32_t)(((x % (int)X) + (int)X) % (int)X);
    return ((row[xw >> 3] >> (4u * (xw & 7u))) & 15u) == M ? 1u : 0u;
}

// WRAP: every axis wraps (rayMarchCompute.slang:80-108 with WRAP_AT_BOUNDARY
// on one domain); each of the 26 offsets is its own term, so on an axis
// shorter than 3 cells the same cell -- or the cell itself -- is counted once
// per offset that reaches it.
template <bool WRAP>
__global__ void __launch_bounds__(256) k_step_gather(const uint32_t *__restrict__ in, uint32_t *__restrict__ out,
                                                     GridGeom g, DevRule rule, StatPart *__restrict__ partials,
                                                     const uint32_t *__restrict__ done)
{
    __shared__ int32_t red[10 * 4];
    if (done && *done) return;  // the manager already ended this simulation
    StatAcc acc;
    const size_t n = (size_t)g.pitch * g.Y * g.Z;
    const uint32_t upr = (g.X + 7u) / 8u;
    // grid-stride over the words: the launch is capped at kMaxStepBlocks, so
    // the per-tick partials stay a few hundred KiB at any grid size
    for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < n; w += (size_t)gridDim.x * blockDim.x) {
        const uint32_t xw = (uint32_t)(w % g.pitch);
        const size_t row = w / g.pitch;
        const int y = (int)(row % g.Y);
        const int z = (int)(row / g.Y);
        if (xw < upr) {
            // 24-bit live masks of the 9 rows around (y, z): [word-1 | word | word+1]
            uint32_t rm[3][3];
#pragma unroll
            for (int dz = -1; dz <= 1; dz++)
#pragma unroll
                for (int dy = -1; dy <= 1; dy++) {
                    int yy = y + dy, zz = z + dz;
                    uint32_t m = 0;
                    if constexpr (WRAP) {
                        // bits 7..16: cells 8*xw-1 .. 8*xw+8, each taken modulo X (a
                        // cell past the row end is the row's first cells)
                        yy = (yy + (int)g.Y) % (int)g.Y;
                        zz = (zz + (int)g.Z) % (int)g.Z;
                        const uint32_t *rowp = in + ((size_t)zz * g.Y + yy) * g.pitch;
                        for (int j = 0; j < 10; j++)
                            m |= live_bit_wrapped(rowp, (int)(xw * 8u) - 1 + j, g.X, rule.M) << (7 + j);
                    } else if (yy >= 0 && yy < (int)g.Y && zz >= 0 && zz < (int)g.Z) {
                        const size_t base = ((size_t)zz * g.Y + yy) * g.pitch;
                        const uint32_t wl = xw > 0 ? in[base + xw - 1] : 0u;
                        const uint32_t wc = in[base + xw];
                        const uint32_t wr = xw + 1 < upr ? in[base + xw + 1] : 0u;
                        m = live_mask8(wl, rule.M) | (live_mask8(wc, rule.M) << 8) | (live_mask8(wr, rule.M) << 16);
                    }
                    rm[dz + 1][dy + 1] = m;
                }
            const uint32_t cur_word = in[w];
            uint32_t out_word = 0;
            const bool fF = rule.flags & 1u, fE = rule.flags & 2u, fC = rule.flags & 4u;
#pragma unroll
            for (uint32_t i = 0; i < 8; i++) {
                const uint32_t x = xw * 8u + i;
                if (x >= g.X) break;
                uint32_t nb = 0;
#pragma unroll
                for (int dz = -1; dz <= 1; dz++)
#pragma unroll
                    for (int dy = -1; dy <= 1; dy++)
#pragma unroll
                        for (int dx = -1; dx <= 1; dx++) {
                            const int cls = (dx != 0) + (dy != 0) + (dz != 0);
                            const bool en = cls == 1 ? fF : cls == 2 ? fE : cls == 3 ? fC : false;
                            if (en) nb += (rm[dz + 1][dy + 1] >> (8 + (int)i + dx)) & 1u;
                        }
                const uint32_t cur = (cur_word >> (4u * i)) & 15u;
                uint32_t ns;
                // searchSimulationCube.slang:178-211
                if (cur == rule.M) {
                    if ((rule.surv_lut >> nb) & 1u) { ns = rule.M; acc.surv++; }
                    else { ns = cur - 1u; acc.dying++; }
                } else if (cur == 0u) {
                    if ((rule.born_lut >> nb) & 1u) { ns = rule.M; acc.born++; }
                    else { ns = 0u; acc.dead++; }
                } else {
                    ns = cur - 1u;
                    acc.dying++;
                }
                out_word |= (ns & 15u) << (4u * i);
                if (ns == rule.M) {
                    acc.bmin[0] = min(acc.bmin[0], (int32_t)x); acc.bmax[0] = max(acc.bmax[0], (int32_t)x);
                    acc.bmin[1] = min(acc.bmin[1], (int32_t)y); acc.bmax[1] = max(acc.bmax[1], (int32_t)y);
                    acc.bmin[2] = min(acc.bmin[2], (int32_t)z); acc.bmax[2] = max(acc.bmax[2], (int32_t)z);
                }
            }
            out[w] = out_word;
        } else {
            out[w] = 0u;
        }
    }
    block_reduce_stats(acc, &partials[blockIdx.x], red);
}

// ---------------------------------------------------------------------------
// per-tick reduction of block partials: one block per tick
// ---------------------------------------------------------------------------
// out may be host memory mapped into the device (the step handle's staging).
// With `flag` set, every block counts itself done on `count` (device memory,
// 0 on entry) after its row is written; the last block resets the count and,
// once every row is visible to the host, writes 1 to the host-mapped flag,
// which the host polls instead of waiting for the stream.
__global__ void __launch_bounds__(256) k_reduce_ticks(const StatPart *__restrict__ partials, uint32_t nparts,
                                                      CaveTickStats *__restrict__ out, uint32_t *count,
                                                      uint32_t *flag)
{
    __shared__ int32_t red[10 * 4];
    const StatPart *p = partials + (size_t)blockIdx.x * nparts;
    StatAcc acc;
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
        const StatPart q = p[i];
        acc.dead += q.dead; acc.dying += q.dying; acc.born += q.born; acc.surv += q.surv;
        for (int k = 0; k < 3; k++) {
            acc.bmin[k] = min(acc.bmin[k], q.bmin[k]);
            acc.bmax[k] = max(acc.bmax[k], q.bmax[k]);
        }
    }
    __shared__ StatPart tot;
    block_reduce_stats(acc, &tot, red);
    __syncthreads();
    if (threadIdx.x == 0) {
        CaveTickStats s;
        s.snap.numberOfDeadCells = tot.dead;
        s.snap.numberOfDyingCells = tot.dying;
        s.snap.numberOfCellsBorn = tot.born;
        s.snap.numberOfSurvivingCells = tot.surv;
        s.live = tot.born + tot.surv;
        s.changed = tot.born + tot.dying;
        for (int k = 0; k < 3; k++) {
            // empty box reported as min = 0x7fffffff, max = -1
            s.bbox_min[k] = tot.bmin[k];
            s.bbox_max[k] = tot.bmax[k];
        }
        out[blockIdx.x] = s;
        if (flag) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");  // the row reaches memory the host reads
            const uint32_t prev = __hip_atomic_fetch_add(count, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
            if (prev + 1u == gridDim.x) {
                __hip_atomic_store(count, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "");
                __hip_atomic_store(flag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }
}

// ---------------------------------------------------------------------------
// plain per-config manager (one block): reduce the tick's block partials and
// apply searchManager.slang:115-139 on the device
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_plain_manager(const StatPart *__restrict__ partials, uint32_t nparts,
                                                       PlainState *__restrict__ s, uint32_t ticks,
                                                       uint32_t sim_flags, CaveTickStats *__restrict__ per_tick,
                                                       uint32_t total_cells)
{
    __shared__ int32_t red[10 * 4];
    __shared__ StatPart tot;
    if (s->done) return;
    StatAcc acc;
    for (uint32_t i = threadIdx.x; i < nparts; i += blockDim.x) {
        const StatPart q = partials[i];
        acc.dead += q.dead; acc.dying += q.dying; acc.born += q.born; acc.surv += q.surv;
        for (int k = 0; k < 3; k++) {
            acc.bmin[k] = min(acc.bmin[k], q.bmin[k]);
            acc.bmax[k] = max(acc.bmax[k], q.bmax[k]);
        }
    }
    block_reduce_stats(acc, &tot, red);
    __syncthreads();
    if (threadIdx.x == 0) {
        CaveGridInfo info = s->info;
        info.numberOfSimulationTicksSurvived++;
        info.totalNumberOfCellsThatDied += tot.dead;
        info.totalNumberOfCellsThatWereBorn += tot.born;
        info.finalAliveCount = tot.surv;
        info.maxAliveCount = max(info.maxAliveCount, tot.surv);
        const bool died = (sim_flags & CAVE_SIM_LEGACY_DEATH) ? tot.dead == total_cells
                                                              : (tot.surv == 0u && tot.born == 0u);
        if (died) info.status = 1u;
        else if (info.numberOfSimulationTicksSurvived >= ticks) info.status = 2u;
        else info.status = 0u;
        CaveTickStats l;
        l.snap.numberOfDeadCells = tot.dead;
        l.snap.numberOfDyingCells = tot.dying;
        l.snap.numberOfCellsBorn = tot.born;
        l.snap.numberOfSurvivingCells = tot.surv;
        l.live = tot.born + tot.surv;
        l.changed = tot.born + tot.dying;
        for (int k = 0; k < 3; k++) {
            l.bbox_min[k] = tot.bmin[k];
            l.bbox_max[k] = tot.bmax[k];
        }
        if (per_tick) per_tick[info.numberOfSimulationTicksSurvived - 1u] = l;
        s->info = info;
        s->last = l;
        s->done = ((info.status == 1u && !(sim_flags & CAVE_SIM_NO_EARLY_EXIT)) ||
                   info.numberOfSimulationTicksSurvived >= ticks) ? 1u : 0u;
    }
}

int launch_plain_manager(const StatPart *partials, uint32_t nparts, PlainState *state, uint32_t ticks,
                         uint32_t sim_flags, CaveTickStats *per_tick, uint32_t total_cells, hipStream_t st)
{
    hipLaunchKernelGGL(k_plain_manager, dim3(1), dim3(256), 0, st, partials, nparts, state, ticks, sim_flags, per_tick,
                       total_cells);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

int launch_init(uint32_t *cells, const GridGeom &g, const DevInit &init, uint32_t M, hipStream_t st)
{
    if (g.planar) return launch_init_planar(cells, g, init, M, st);
    const size_t n = (size_t)g.pitch * g.Y * g.Z;
    const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 65535u * 4u);
    hipLaunchKernelGGL(k_init_cells, dim3(blocks), dim3(256), 0, st, cells, g, init, M & 15u);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

// Zeroes every nibble at x >= X (row padding) of a row-pitched nibble layout:
// the reference writes padding as 0 (searchSimulationCube.slang:161-166) and
// never reads it as a neighbour, so caller-supplied cells are normalised once.
__global__ void k_sanitize_nibbles(uint32_t *cells, GridGeom g)
{
    const size_t n = (size_t)g.pitch * g.Y * g.Z;
    for (size_t w = blockIdx.x * (size_t)blockDim.x + threadIdx.x; w < n; w += (size_t)gridDim.x * blockDim.x) {
        const uint32_t x0 = (uint32_t)(w % g.pitch) * 8u;
        if (x0 + 8u <= g.X) continue;
        const uint32_t keep = x0 >= g.X ? 0u : (0xFFFFFFFFu >> (4u * (8u - (g.X - x0))));
        cells[w] &= keep;
    }
}

int launch_sanitize_nibbles(uint32_t *cells, const GridGeom &g, hipStream_t st)
{
    const size_t n = (size_t)g.pitch * g.Y * g.Z;
    const uint32_t blocks = (uint32_t)std::min<size_t>((n + 255) / 256, 65535u * 4u);
    hipLaunchKernelGGL(k_sanitize_nibbles, dim3(blocks), dim3(256), 0, st, cells, g);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

uint32_t gather_blocks(const GridGeom &g)
{
    const size_t n = (size_t)g.pitch * g.Y * g.Z;
    return (uint32_t)std::min<size_t>((n + 255) / 256, kMaxStepBlocks);
}

int launch_gather(const uint32_t *in, uint32_t *out, const GridGeom &g, const DevRule &rule, StatPart *partials,
                  hipStream_t st, const uint32_t *done, bool wrap)
{
    if (wrap)
        hipLaunchKernelGGL(k_step_gather<true>, dim3(gather_blocks(g)), dim3(256), 0, st, in, out, g, rule, partials,
                           done);
    else
        hipLaunchKernelGGL(k_step_gather<false>, dim3(gather_blocks(g)), dim3(256), 0, st, in, out, g, rule, partials,
                           done);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

int launch_reduce_ticks(const StatPart *partials, uint32_t nparts, uint32_t ticks, CaveTickStats *out,
                        hipStream_t st, uint32_t *count, uint32_t *flag)
{
    if (ticks == 0) return CAVE_OK;
    hipLaunchKernelGGL(k_reduce_ticks, dim3(ticks), dim3(256), 0, st, partials, nparts, out, count, flag);
    CAVE_HIP(hipGetLastError());
    return CAVE_OK;
}

}  // namespace cave
