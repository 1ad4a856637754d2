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

// ---- wavefront reductions (64 lanes) via DPP-backed cross-lane ops ----
__device__ inline uint32_t wave_sum_u32(uint32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ inline int32_t wave_min_i32(int32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = min(v, __shfl_xor(v, off, 64));
    return v;
}

__device__ inline int32_t wave_max_i32(int32_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
    return v;
}

// Thread-local statistics accumulator.
struct StatAcc {
    uint32_t dead = 0, dying = 0, born = 0, surv = 0;
    int32_t bmin[3] = {0x7fffffff, 0x7fffffff, 0x7fffffff};
    int32_t bmax[3] = {-1, -1, -1};
};

// Reduces a StatAcc over the block; thread 0 returns the block total in *out.
// scratch must hold 10 * (blockDim.x / 64) words of LDS.
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
    if (threadIdx.x == 0) {
        StatPart p;
        p.dead = p.dying = p.born = p.surv = 0;
        p.bmin[0] = p.bmin[1] = p.bmin[2] = 0x7fffffff;
        p.bmax[0] = p.bmax[1] = p.bmax[2] = -1;
        for (int w = 0; w < nwaves; w++) {
            const int32_t *v = scratch + w * 10;
            p.dead += (uint32_t)v[0]; p.dying += (uint32_t)v[1];
            p.born += (uint32_t)v[2]; p.surv += (uint32_t)v[3];
            for (int k = 0; k < 3; k++) {
                p.bmin[k] = min(p.bmin[k], v[4 + k]);
                p.bmax[k] = max(p.bmax[k], v[7 + k]);
            }
        }
        *out = p;
    }
}

}  // namespace cave
