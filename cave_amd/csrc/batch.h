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

// Words of cell storage a job needs: 4 state planes + 2 live planes of
// ceil(X/32) words per row.
inline uint64_t job_words(uint32_t X, uint32_t Y, uint32_t Z)
{
    return 6ull * ((X + 31u) / 32u) * (uint64_t)Y * Z;
}

int launch_batch(const BatchArgs &a, uint32_t blocks, uint32_t threads, uint32_t lds_bytes, hipStream_t st);

}  // namespace cave
