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
    uint32_t tier;         // kTierLds, kTierBlockArena or kTierOwnRegion
    uint32_t result_slot;  // index into results / per-tick rows / completion flags
    uint32_t has_final;    // 1: store the final cells (reference packed layout) at final_addr
    uint32_t zchunks;      // z-chunks per column per tick over all parts (a multiple of parts)
    uint32_t parts;        // blocks that share the job (1 = one block owns it)
    uint32_t part;         // this queue entry's part: z-chunks [part*c, (part+1)*c), c = zchunks/parts
    uint32_t sync_slot;    // JobSync record of a multi-part job
    uint32_t pad_;
    uint64_t region;       // device address of the job's own cells (kTierOwnRegion)
    uint64_t final_addr;   // device address of its final cells (has_final)
    uint64_t seq;          // served requests: completion value written to BatchArgs::done[result_slot]
    DevInit init;
};

// Where a job's cells live while it runs.
constexpr uint32_t kTierLds = 0;         // the block's LDS, for all of its ticks
constexpr uint32_t kTierBlockArena = 1;  // the claiming block's private region of the global arena
constexpr uint32_t kTierOwnRegion = 2;   // a region of the job's own (multi-part jobs, served requests)

// Cross-block rendezvous of the parts of one job.  Tick t uses counter slot
// t % 3: parts add their block partials, arrive, wait for all parts, then read
// the job totals.  Zeroed before the launch.
struct JobSync {
    uint32_t arrive;        // arrivals over all phases (init = phase 0, tick t = phase t + 1)
    uint32_t abort;         // set when a wait timed out: every part leaves the job
    uint32_t pad_[2];
    uint32_t sum[3][4];     // dying, born, surviving
    int32_t nmin[3][4];     // 0x7fffffff - bbox min (max-combined, 0 = empty)
    int32_t xmax1[3][4];    // bbox max + 1 (max-combined, 0 = empty)
    // per-tick rendezvous of jobs whose totals fit a word: surviving (bits
    // 0-23) | born (24-47) | arrivals (48-63), one slot per tick mod 3
    uint64_t tick[3];
    uint64_t pad2_;
    // those jobs' per-tick bounding boxes, read one tick later: slot tick % 4,
    // encoded as nmin / xmax1
    int32_t lag_nmin[4][4];
    int32_t lag_xmax1[4][4];
};

// Result record written by the kernel (host converts to CaveSimResult).
struct DevResult {
    CaveGridInfo info;
    uint32_t classification;
    uint32_t ticks_run;
    CaveTickStats last;
    uint64_t t_start, t_end;  // device wall clock at claim / at completion
    uint32_t err;             // 1: a part rendezvous timed out
    uint32_t pad_;
    uint64_t computed;        // cells of the active regions of its ticks (every part's)
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
    uint32_t lds_words;          // dynamic LDS words available
    JobSync *sync;               // one record per multi-part job
    uint64_t wait_limit;         // device wall-clock ticks a part waits for the others
    // served requests only (k_serve)
    const struct ServeCtl *ctl;  // host-coherent: ring tail, stop flag
    struct ServeDev *sd;         // device-side claim state
    uint64_t ring_cap;           // ring entries (jobs[] is the host-coherent ring)
    uint64_t *done;              // host-coherent completion flags, [result_slot]
    uint64_t idle_limit;         // device wall-clock ticks without a claim before the kernel closes
};

// Control block of the request ring, in host-coherent memory: the host writes
// ring entries, then publishes them by raising `tail`; `stop` asks every block
// to exit once it has no work.
struct ServeCtl {
    uint64_t tail;
    uint64_t stop;
    uint64_t pad_[6];
};

// Claim state of the resident kernel, in device memory (zeroed before every
// launch except `head`, which carries over between launches).
struct ServeDev {
    uint64_t head;        // ring entries claimed so far
    uint64_t tail_seen;   // latest ring tail read from the host
    uint64_t last_claim;  // device clock of the latest claim
    uint32_t poll_lock;   // held by the one block reading the host's tail
    uint32_t closing;     // set once: every block leaves the claim loop
};

// Words of cell storage a job needs: 4 state planes of ceil(X/32) words per
// row, and 2 live planes with a one-row / one-plane zero guard band, a guard
// word after each row and one before the plane.
inline uint64_t job_words(uint32_t X, uint32_t Y, uint32_t Z)
{
    const uint64_t W = (X + 31u) / 32u;
    return 4ull * W * Y * Z + 2ull * ((W + 1ull) * (Y + 2ull) * (Z + 2ull) + 1ull);
}

// Work split of one tick: columns (x-word, y) times z-chunks, about `ipt`
// items per thread of the block when the grid allows it.  With more columns
// than threads, the items run in rounds, and an item of c planes takes c + 2
// steps (two halo planes).  There, the z-chunk count minimises rounds x
// (c + 2): a 6 x 191 x 48 part on 1024 threads takes 4 chunks (5 rounds of
// 14 steps) rather than 1 (2 rounds of 50, the second with 122 items).
inline uint32_t job_zchunks(uint32_t X, uint32_t Y, uint32_t Z, uint32_t threads, uint32_t ipt)
{
    const uint32_t cols = ((X + 31u) / 32u) * Y;
    const uint32_t T = threads ? threads : 1u;
    if (cols > T) {
        uint32_t best = 1u;
        uint64_t best_cost = ~0ull;
        for (uint32_t z = 1u; z <= Z && z <= 64u; z++) {
            const uint32_t c = (Z + z - 1u) / z;
            const uint64_t used = (Z + c - 1u) / c;
            const uint64_t rounds = ((uint64_t)cols * used + T - 1u) / T;
            const uint64_t cost = rounds * (c + 2u);
            if (cost < best_cost) best = z, best_cost = cost;
        }
        return best;
    }
    uint32_t z = T * (ipt ? ipt : 1u) / (cols ? cols : 1u);
    if (z < 1u) z = 1u;
    if (z > Z) z = Z;
    return z;
}

// z-chunks of one part of a job split over `parts` blocks, as one job of its
// planes, times the number of parts.
inline uint32_t part_zchunks(uint32_t X, uint32_t Y, uint32_t Z, uint32_t threads, uint32_t ipt, uint32_t parts)
{
    const uint32_t zp = (Z + parts - 1u) / parts;
    return job_zchunks(X, Y, zp, threads, ipt) * parts;
}

// Cell shape of a persistent launch (every job of a launch has the same):
// the cube (W ignored), the ERD shape, the ERD shape with every axis wrapping.
enum : uint32_t { kBatchShapeCube = 0, kBatchShapeErd = 1, kBatchShapeErdWrap = 2 };
const void *batch_kernel(bool split, uint32_t shape);
const void *serve_kernel();
int launch_serve(const BatchArgs &a, uint32_t blocks, uint32_t threads, uint32_t lds_bytes, hipStream_t st);
// split_jobs: the batch holds jobs split over several blocks (selects the
// kernel form with the one-word per-tick rendezvous).
int launch_batch(const BatchArgs &a, uint32_t blocks, uint32_t threads, uint32_t lds_bytes, hipStream_t st,
                 bool split_jobs, uint32_t shape);

}  // namespace cave
