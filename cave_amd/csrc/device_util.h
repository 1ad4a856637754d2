// device_util.h -- wavefront / block reductions and the cell-init hash.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"

namespace cave {

// ---- Thomas Wang 32-bit integer hash (searchSystem.cpp:244-259 / cellInit.slang:33-42) ----
__host__ __device__ inline uint32_t wang_hash(uint32_t s)
{
    s = (s ^ 61u) ^ (s >> 16);
    s *= 9u;
    s = s ^ (s >> 4);
    s *= 0x27d4eb2du;
    s = s ^ (s >> 15);
    return s;
}

__host__ __device__ inline uint32_t cell_hash(uint32_t x, uint32_t y, uint32_t z, uint32_t seed)
{
    uint32_t h = wang_hash(x * 73856093u + seed);
    h = wang_hash(h ^ (y * 19349663u));
    h = wang_hash(h ^ (z * 83492791u));
    return h;
}

// Init descriptor shared by every kernel that builds initial cells in place.
struct DevInit {
    uint32_t box_lo[3];   // spawn box [lo, hi) per axis
    uint32_t box_hi[3];
    uint32_t random;      // 1: hash < threshold (unless all_in), 0: filled
    uint32_t all_in;      // density >= 1
    uint32_t threshold;
    uint32_t seed;
};

__host__ __device__ inline bool init_alive(const DevInit &in, uint32_t x, uint32_t y, uint32_t z)
{
    bool box = x >= in.box_lo[0] && x < in.box_hi[0] && y >= in.box_lo[1] && y < in.box_hi[1] &&
               z >= in.box_lo[2] && z < in.box_hi[2];
    if (!box) return false;
    if (!in.random || in.all_in) return true;
    return cell_hash(x, y, z, in.seed) < in.threshold;
}

// ---- wavefront reductions (64 full lanes) with DPP ----
// No LDS traffic: two quad swaps and two 16-lane row rotations reduce each
// row, then row broadcasts of lanes 15 and 31 fold the rows into lane 63,
// whose value is returned to every lane.  Lanes a row mask leaves out take
// `ident`, the operation's identity.
template <typename Op>
__device__ __forceinline__ int32_t wave_reduce_dpp(int32_t v, int32_t ident, Op op)
{
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x124, 0xF, 0xF, false));  // row_ror:4
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x128, 0xF, 0xF, false));  // row_ror:8
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x142, 0xA, 0xF, false));  // row_bcast:15 into rows 1, 3
    v = op(v, __builtin_amdgcn_update_dpp(ident, v, 0x143, 0xC, 0xF, false));  // row_bcast:31 into rows 2, 3
    return __builtin_amdgcn_readlane(v, 63);
}

__device__ inline uint32_t wave_sum_u32(uint32_t v)
{
    return (uint32_t)wave_reduce_dpp((int32_t)v, 0, [](int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); });
}

__device__ inline int32_t wave_min_i32(int32_t v)
{
    return wave_reduce_dpp(v, 0x7fffffff, [](int32_t a, int32_t b) { return min(a, b); });
}

__device__ inline int32_t wave_max_i32(int32_t v)
{
    return wave_reduce_dpp(v, (int32_t)0x80000000, [](int32_t a, int32_t b) { return max(a, b); });
}

// Thread-local statistics accumulator.
struct StatAcc {
    uint32_t dead = 0, dying = 0, born = 0, surv = 0;
    int32_t bmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
    int32_t bmax[3] = {-1, -1, -1};
};

// Reduces a StatAcc over the block (at most 1024 threads); thread 0 returns
// the block total in *out.  scratch must hold 10 * (blockDim.x / 64) words of
// LDS.
__device__ inline void block_reduce_stats(const StatAcc &a, StatPart *out, int32_t *scratch)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = (blockDim.x + 63) >> 6;
    uint32_t s0 = wave_sum_u32(a.dead), s1 = wave_sum_u32(a.dying);
    uint32_t s2 = wave_sum_u32(a.born), s3 = wave_sum_u32(a.surv);
    int32_t m0 = wave_min_i32(a.bmin[0]), m1 = wave_min_i32(a.bmin[1]), m2 = wave_min_i32(a.bmin[2]);
    int32_t x0 = wave_max_i32(a.bmax[0]), x1 = wave_max_i32(a.bmax[1]), x2 = wave_max_i32(a.bmax[2]);
    if (lane == 0) {
        int32_t *w = scratch + wave * 10;
        w[0] = (int32_t)s0; w[1] = (int32_t)s1; w[2] = (int32_t)s2; w[3] = (int32_t)s3;
        w[4] = m0; w[5] = m1; w[6] = m2; w[7] = x0; w[8] = x1; w[9] = x2;
    }
    __syncthreads();
    if (wave == 0) {  // lane w < 16 reduces wave w's partials (nwaves <= 16: one DPP row)
        // Every lane loads (lanes past nwaves read wave 0's slot and drop it),
        // so the ten loads issue together behind one wait; the ten reductions
        // then take the four in-row DPP steps side by side.
        const bool has = lane < nwaves;
        const int32_t *v = scratch + (has ? lane : 0) * 10;
        int32_t x[10];
        for (int k = 0; k < 10; k++) x[k] = v[k];
        for (int k = 0; k < 4; k++) x[k] = has ? x[k] : 0;
        for (int k = 4; k < 7; k++) x[k] = has ? x[k] : 0x7fffffff;
        for (int k = 7; k < 10; k++) x[k] = has ? x[k] : -1;
#define CAVE_ROW16_STEP(ctl)                                                                              \
        for (int k = 0; k < 4; k++)                                                                      \
            x[k] = (int32_t)((uint32_t)x[k] + (uint32_t)__builtin_amdgcn_update_dpp(0, x[k], ctl, 0xF, 0xF, false)); \
        for (int k = 4; k < 7; k++) x[k] = min(x[k], __builtin_amdgcn_update_dpp(0x7fffffff, x[k], ctl, 0xF, 0xF, false)); \
        for (int k = 7; k < 10; k++) x[k] = max(x[k], __builtin_amdgcn_update_dpp(-1, x[k], ctl, 0xF, 0xF, false));
        CAVE_ROW16_STEP(0xB1)   // quad_perm [1,0,3,2]
        CAVE_ROW16_STEP(0x4E)   // quad_perm [2,3,0,1]
        CAVE_ROW16_STEP(0x124)  // row_ror:4
        CAVE_ROW16_STEP(0x128)  // row_ror:8
#undef CAVE_ROW16_STEP
        if (lane == 0) {
            StatPart p;
            p.dead = (uint32_t)x[0];
            p.dying = (uint32_t)x[1];
            p.born = (uint32_t)x[2];
            p.surv = (uint32_t)x[3];
            for (int k = 0; k < 3; k++) {
                p.bmin[k] = x[4 + k];
                p.bmax[k] = x[7 + k];
            }
            *out = p;
        }
    }
}

}  // namespace cave
