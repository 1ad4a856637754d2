// batch_plan.h -- how a list of simulations is laid out on the persistent
// batched kernel: block size, storage tier per simulation, z-chunk tiling,
// splitting of long simulations over several blocks, and queue order.  The
// choices come from a rate table that the autotuner fills by timing short
// calibration launches per (grid-edge bucket, neighbourhood) bucket.
#pragma once

#include <cstdint>
#include <vector>

#include "batch.h"

namespace cave {

struct BatchKnobs {
    uint32_t threads = 0;    // block size (0 = chosen by the tuner, 1024 without it)
    uint32_t max_parts = 0;  // most blocks per simulation (0 = auto, 1 = never split)
    uint32_t ipt = 0;        // work items per thread of a tick (0 = per bucket from the tuner, else fixed)
    bool tune = true;        // consult / fill the rate table
};

struct BatchPlan {
    uint32_t threads = 1024, blocks = 1, lds_bytes = 0;
    std::vector<DevJob> jobs;   // one per spec, in spec order
    std::vector<DevJob> queue;  // one entry per part, in claim order
    std::vector<uint32_t> ipt;  // work items per thread, per spec
    uint64_t arena_words = 0;   // per block (kTierBlockArena)
    uint64_t region_words = 0;  // total of kTierOwnRegion regions
    uint32_t nsync = 0;
    uint64_t final_words = 0;   // staging for final cells (CAVE_SIM_FINAL_HASH)
    double predicted_ms = 0;    // the tuner's estimate of the launch (0 without tuning)
    uint32_t shape = kBatchShapeCube;  // the jobs' cell shape (a plan holds one shape)
};

// The rate table's neighbourhood key: F|E|C for cube jobs; for the ERD
// shape, whose ticks cost differently, kTuneErd | F|E, plus 4 with wrap.
constexpr uint32_t kTuneErd = 16u;
// Batch-plan keys only (the serve path's keys never carry them): jobs that
// compute every cell (CAVE_SIM_FULL_GRID), and rates measured on a sample of
// the list's own specs rather than on the synthetic grid.
constexpr uint32_t kTuneFull = 32u;
constexpr uint32_t kTuneSample = 64u;
uint32_t tune_kind(const DevJob &j, uint32_t shape);
// The persistent kernel form (kBatchShape*) a spec runs on.
uint32_t spec_shape(const CaveSimSpec &s);

// Grid-edge bucket of a simulation (edge = cube root of its cell count):
// [0,32) [32,48) [48,64) [64,96) [96,128) [128,...).
uint32_t edge_bucket(uint32_t X, uint32_t Y, uint32_t Z);

// Blocks of `threads` threads one CU keeps resident with `lds_bytes` of
// dynamic LDS each.
uint32_t resident_per_cu(uint32_t threads, uint32_t lds_bytes, uint32_t shape);
uint32_t cu_count();
DevJob make_job(const CaveSimSpec &s, uint32_t slot);

int plan_batch(const CaveSimSpec *specs, size_t n, const BatchKnobs &k, BatchPlan &plan);

// LDS a block of `threads` threads may use so that 1024 / threads blocks share a CU.
uint32_t lds_cap(uint32_t threads);

// Device-wide rate (cell-updates/s, every resident block on the bucket) of a
// (grid-edge bucket, neighbourhood) at a block size and items per thread from
// the tuner's table.  When missing and `measure` is set, one calibration
// launch on `st` with a grid of `dims` measures it; 0 when unknown.
double bucket_rate(uint32_t bucket, uint32_t flags, uint32_t threads, uint32_t ipt, const uint32_t dims[3],
                   bool measure, hipStream_t st);

uint64_t wall_clock_khz();

// Runs a planned batch on `st` and fills results (and the options' per-tick
// rows, timing and device time).
int run_plan(const BatchPlan &plan, size_t n, CaveSimResult *results, CaveBatchOptions *opt, hipStream_t st);

}  // namespace cave
