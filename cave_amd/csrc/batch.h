// batch.h -- device job descriptors of the persistent batched engine.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "bitslice.h"
#include "common.h"
#include "device_util.h"

namespace cave {

// One simulation (one rule config on one grid with one seed).
struct DevJob {
    uint32_t X, Y, Z;
    uint32_t flags;        // F|E|C
    uint32_t M;
    uint32_t born_lut, surv_lut;
    uint32_t ticks;        // tick budget
    uint32_t sim_flags;    // CAVE_SIM_*
    uint32_t tier;         // 0: cells resident in LDS, 1: in the global arena
    uint32_t result_slot;  // index into results / per-tick rows
    uint32_t final_slot;   // +1-based index of a final-cells staging area (0 = none)
    uint32_t zchunks;      // z-chunks per column per tick (work items = columns x zchunks)
    DevInit init;
};

// Result record written by the kernel (host converts to CaveSimResult).
struct DevResult {
    CaveGridInfo info;
    uint32_t classification;
    uint32_t ticks_run;
    CaveTickStats last;
};

struct BatchArgs {
    const DevJob *jobs;
    uint32_t njobs;
    uint32_t *next_job;          // work-queue counter (zeroed before launch)
    DevResult *results;
    CaveTickStats *per_tick;     // nullable: [result_slot * per_tick_stride + t]
    uint32_t per_tick_stride;
    uint32_t *arena;             // global-tier cells: one region per block
    uint64_t arena_words;        // words per block region
    uint32_t *final_cells;       // staging for final states (reference layout)
    const uint64_t *final_offsets; // word offset per final slot
    uint32_t lds_words;          // dynamic LDS words available
};

// Words of cell storage a job needs: 4 state planes of ceil(X/32) words per
// row, 2 live planes with a one-row / one-plane zero guard band, and one
// padding word past the end (read, masked, by the last x-neighbour load).
inline uint64_t job_words(uint32_t X, uint32_t Y, uint32_t Z)
{
    const uint64_t W = (X + 31u) / 32u;
    return 4ull * W * Y * Z + 2ull * W * (Y + 2ull) * (Z + 2ull) + 2ull;
}

// Work split of one tick: columns (x-word, y) times z-chunks, about one item
// per thread of the block when the grid allows it.
inline uint32_t job_zchunks(uint32_t X, uint32_t Y, uint32_t Z, uint32_t threads)
{
    const uint32_t cols = ((X + 31u) / 32u) * Y;
    uint32_t z = threads / (cols ? cols : 1u);
    if (z < 1u) z = 1u;
    if (z > Z) z = Z;
    return z;
}

int launch_batch(const BatchArgs &a, uint32_t blocks, uint32_t threads, uint32_t lds_bytes, hipStream_t st);

}  // namespace cave
