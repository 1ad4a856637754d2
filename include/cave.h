/*
 * cave.h -- C-ABI of the MI355X-native CAVE step + rule-search engine.
 *
 * Plain C linkage, plain pointers and sizes.  Host buffers are caller-owned;
 * device memory is library-owned.  Every entry point returns an int status
 * (CAVE_OK = 0) and never throws; cave_last_error() gives the message of the
 * calling thread's last failure.  A handle is used by one thread at a time;
 * different handles may be used from different threads.
 *
 * Reference interfaces replaced (paths relative to the CAVE repository):
 *   CaveSimulationParameters  <- SimulationParameters   src/common/structs.h:89-99
 *   CaveGridSnapshot          <- GridSnapshot           src/common/structs.h:130-136
 *   CaveGridInfo              <- GridInfo               src/common/structs.h:79-87
 *   CaveChunkConfiguration    <- ChunkConfiguration     src/common/structs.h:109-128
 *   cave_step_*               <- ComputeSystem::Tick + the headless step loop
 *                                src/vulkan/systems/computeSystem.h:54-56,
 *                                src/modes/renderingMode.cpp:166-343, with the
 *                                search-kernel transition of
 *                                shaders/slang/search/searchSimulationCube.slang:78-247
 *   cave_search_*             <- SearchSystem (ctor, IsTickComplete/FlushProgress/
 *                                SubmitNextTick, IsRunning, GetProgress,
 *                                Get*Results, WriteToDisk)
 *                                src/vulkan/systems/searchSystem.h:218-275
 *   cave_rule_string_to_bitmask / cave_bitmask_to_rule_string
 *                             <- GuiSystem::RuleStringToBitmask / BitmaskToRuleString
 *                                src/vulkan/systems/guiSystem.cpp:92-163
 *   cave_enumerate_valid_bitmasks
 *                             <- PermutationMath::EnumerateValidBitmasks
 *                                src/common/permutationMath.h:101-128
 *   cave_batch_*, cave_serve_* have no reference counterpart: batched
 *   evaluation of explicit (rule, C, grid, seed, ticks) lists and a
 *   continuous-batching request server.
 */
#ifndef CAVE_H
#define CAVE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CAVE_ABI_VERSION 2

/* ---- status codes ---- */
enum {
    CAVE_OK = 0,
    CAVE_E_INVALID = 1,     /* bad argument / spec                      */
    CAVE_E_UNSUPPORTED = 2, /* valid in the reference, not in the form asked for here */
    CAVE_E_NOMEM = 3,       /* device or host allocation failed         */
    CAVE_E_DEVICE = 4,      /* HIP runtime error                        */
    CAVE_E_BUDGET = 5,      /* request larger than the admission budget */
    CAVE_E_STATE = 6        /* call not valid in the handle's state     */
};

/* ---- neighbourhood flags (src/simulation/simulation.h:47-50) ---- */
#define CAVE_FACES 1u
#define CAVE_EDGES 2u
#define CAVE_CORNERS 4u
#define CAVE_WRAP 8u

/* ---- reference structs, byte-identical ---- */
typedef struct CaveSimulationParameters {
    uint64_t birthAndMaxCellStateRules;    /* bits 60-63: M-1; bits 0-59: birth mask or 1-based index */
    uint64_t survivalAndNeighborhoodRules; /* bits 60-63: F|E|C|W flags; bits 0-59: survival mask or index */
    uint64_t shapeAndGridConfiguration;    /* bits 0-4: shape id (0 = cube) */
} CaveSimulationParameters;

typedef struct CaveGridSnapshot {
    uint32_t numberOfDeadCells;      /* 0 -> 0                 */
    uint32_t numberOfDyingCells;     /* M -> M-1 or 0<s<M -> s-1 */
    uint32_t numberOfCellsBorn;      /* 0 -> M                 */
    uint32_t numberOfSurvivingCells; /* M -> M                 */
} CaveGridSnapshot;

typedef struct CaveGridInfo {
    uint32_t totalNumberOfCellsThatDied;
    uint32_t totalNumberOfCellsThatWereBorn;
    uint32_t numberOfSimulationTicksSurvived;
    uint32_t status; /* 0 running, 1 died, 2 reached the tick budget */
    uint32_t maxAliveCount;
    uint32_t finalAliveCount;
} CaveGridInfo;

typedef struct CaveChunkConfiguration {
    uint32_t shapeId;
    uint32_t gridDimensionX, gridDimensionY, gridDimensionZ;
    uint32_t spawnAreaDimensionX, spawnAreaDimensionY, spawnAreaDimensionZ;
    uint32_t spawnPlacementId;
    uint32_t spawnRegionShapeId; /* 0 = random, 1 = filled */
    uint32_t spawnSizeIndex;
    uint32_t spawnDensityIndex;  /* density = (index + 1) / 16 */
    uint32_t runSeedIndex;
    uint32_t partitionIndex;
    uint32_t partitionCount;
} CaveChunkConfiguration;

/* Per-tick statistics: the reference counters plus live / changed counts
 * and the bounding box of live (state == M) cells after the tick.
 * bbox_min[a] > bbox_max[a] means no live cell.  The counters are u32, as the
 * reference's are: on grids of 2^32 cells or more they wrap modulo 2^32. */
typedef struct CaveTickStats {
    CaveGridSnapshot snap;
    uint32_t live;    /* born + surviving */
    uint32_t changed; /* born + dying     */
    int32_t bbox_min[3];
    int32_t bbox_max[3];
} CaveTickStats;

/* ---- library ---- */
int cave_abi_version(void);
const char *cave_last_error(void);
/* name of device 0, its CU count and HBM bytes */
int cave_device_info(char *name, size_t cap, int *cu_count, uint64_t *hbm_bytes);

/* ---- host helpers (no device work) ---- */
uint64_t cave_rule_string_to_bitmask(const char *rule);
size_t cave_bitmask_to_rule_string(uint64_t bitmask, char *out, size_t cap);
size_t cave_enumerate_valid_bitmasks(uint32_t neighbor_count, uint32_t max_bits, uint64_t *out, size_t cap);
/* words of the ERD shape's linear layout (cell x + X*(y + Y*z), 8 per word,
 * searchSystem.cpp:296-300): ceil(X*Y*Z/8).  A step handle whose parameters
 * carry shape id 1 (shapeAndGridConfiguration bits 0-4) uploads / downloads
 * this layout. */
size_t cave_erd_cell_words(uint32_t X, uint32_t Y, uint32_t Z);
/* words of the reference packed layout: ceil(X/8) * Y * Z */
size_t cave_cell_words(uint32_t X, uint32_t Y, uint32_t Z);
/* splitmix64 generator used for every synthetic workload (seeded, portable) */
uint64_t cave_splitmix64(uint64_t *state);
/* random rule bitset over counts 1..n_counts, each bit set with probability p */
uint64_t cave_random_rule(uint64_t *state, uint32_t n_counts, double p);

/* ================================================================== */
/* Single-grid step (headless step path)                               */
/* ================================================================== */
typedef struct CaveStep CaveStep;

enum {
    CAVE_SPAWN_RANDOM = 0, /* hash(x,y,z,seed) < floor(density * 2^32) inside the box */
    CAVE_SPAWN_FILLED = 1, /* every box cell live */
    CAVE_SPAWN_BUFFER = 2, /* caller-supplied cells (reference packed layout) */
    /* the render path's initial cells (shaders/slang/simulation/cellInit.slang:33-60):
     * the cell hash without a seed, threshold uint(d * 2^32) computed in float from
     * the "%f" text of the float density (simulationRenderer.cpp:201) */
    CAVE_SPAWN_RENDER_RANDOM = 3
};

enum {
    CAVE_KERNEL_AUTO = 0,
    CAVE_KERNEL_GATHER = 1,        /* one lane per packed word, per-cell gather            */
    CAVE_KERNEL_STREAM = 2,        /* bit-sliced z-streaming kernel, bit-planar cell groups (X <= 2048) */
    CAVE_KERNEL_STREAM_NIBBLE = 3, /* the same kernel on the reference nibble layout        */
    CAVE_KERNEL_ERD = 4            /* the ERD shape (shape id 1), linear layout              */
};

typedef struct CaveStepSpec {
    uint32_t dims[3];
    CaveSimulationParameters params; /* raw-bitmask form (max-rule-bits 0).  The W flag wraps every
                                        axis, as the render path's step does (rayMarchCompute.slang:
                                        84-102); such handles run the planar streaming kernel when
                                        the grid's buffer is under 1 GiB (AUTO or STREAM), else the
                                        gather kernel; STREAM_NIBBLE does not wrap */
    uint32_t spawn[3];               /* spawn box, centred, clamped to dims */
    uint32_t spawn_mode;             /* CAVE_SPAWN_* */
    double density;                  /* random spawn probability, evaluated in double */
    uint32_t seed;                   /* runSeedIndex salt of the cell hash */
    const uint32_t *initial_cells;   /* CAVE_SPAWN_BUFFER: cave_cell_words(dims) words */
    uint32_t kernel;                 /* CAVE_KERNEL_* */
    uint32_t reserved;
} CaveStepSpec;

int cave_step_create(const CaveStepSpec *spec, CaveStep **out);
/* Runs `ticks` ticks.  per_tick (nullable) receives `ticks` entries.
 * device_ms (nullable) receives the device time of the tick loop, measured
 * with HIP events on the handle's stream. */
int cave_step_run(CaveStep *h, uint32_t ticks, CaveTickStats *per_tick, double *device_ms);
/* ---- the render path's cell storage (shaders/slang/common/cellStateUtils.slang:1-70) ----
 * Two cells per byte along x (low nibble = even x).  CAVE_RENDER_IMAGE: the
 * r8ui 3D image, ceil(X/2) x Y x Z bytes, x fastest.  CAVE_RENDER_TILED: the
 * buffer form, one packed byte per u32 element, 8x8x8 texel blocks (block
 * index z-major, then y, then x) with Morton order inside a block
 * (cellStateUtils.slang:18-46).  Host-side conversions to and from the
 * packed search layout (cave_cell_words words). */
enum { CAVE_RENDER_IMAGE = 0, CAVE_RENDER_TILED = 1 };
/* Bytes of a render-layout buffer for a grid. */
size_t cave_render_bytes(uint32_t layout, uint32_t X, uint32_t Y, uint32_t Z);
int cave_render_to_packed(uint32_t layout, const void *render, uint32_t X, uint32_t Y, uint32_t Z,
                          uint32_t *packed);
int cave_packed_to_render(uint32_t layout, const uint32_t *packed, uint32_t X, uint32_t Y, uint32_t Z,
                          void *render);

/* Copies the current cells out in the reference packed layout. */
int cave_step_download(CaveStep *h, uint32_t *cells, size_t n_words);
int cave_step_upload(CaveStep *h, const uint32_t *cells, size_t n_words);
/* Replaces the rule (same M / flags constraints as create). */
int cave_step_set_params(CaveStep *h, const CaveSimulationParameters *params);
/* Which kernel the handle runs (CAVE_KERNEL_GATHER / _STREAM / _STREAM_NIBBLE). */
int cave_step_kernel(const CaveStep *h);
int cave_step_destroy(CaveStep *h);

/* cave_step_run_static: `ticks` ticks of the step on caller-owned device
 * buffers, allocating nothing.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The transition is
 * cave_step_run's (cube shape; the W flag wraps every axis).
 *   cells_a, cells_b: device buffers of cave_cell_words(dims) words in the
 *     reference packed layout; cells_a holds the state before the call, with
 *     row-padding nibbles 0.  Tick t reads one and writes the other, starting
 *     from cells_a, so the state after the call is in cells_a when `ticks` is
 *     even and in cells_b when it is odd.
 *   kernel: CAVE_KERNEL_AUTO (the streaming kernel on the nibble layout when
 *     ceil(X/8) is a multiple of 4, X <= 2048 and no W, else the gather
 *     kernel), CAVE_KERNEL_GATHER or CAVE_KERNEL_STREAM_NIBBLE.
 *   per_tick: nullable; `ticks` CaveTickStats in memory the device can write
 *     (device memory or mapped pinned host memory).
 *   stream: a hipStream_t (NULL: the null stream).  Work is queued on it and
 *     the call returns without waiting.
 * The only scratch is one tick's block partials, in a static device array of
 * 8192 entries (320 KiB per device, held from first use until the process
 * ends); a grid whose plan needs more is rejected with CAVE_E_INVALID.  Every
 * call on a device shares that array, so calls that overlap on one device race. */
int cave_step_run_static(const CaveSimulationParameters *params, const uint32_t dims[3], uint32_t kernel,
                         uint32_t *cells_a, uint32_t *cells_b, uint32_t ticks, CaveTickStats *per_tick,
                         void *stream);

/* ================================================================== */
/* Batched evaluation of explicit configs                              */
/* ================================================================== */
typedef struct CaveSimSpec {
    uint32_t dims[3];
    uint32_t max_cell_state;     /* M, 1..16 (C = M + 1 states) */
    uint32_t neighborhood_flags; /* F|E|C, optionally W: cube simulations ignore W, as the cube
                                    search kernel does (searchSimulationCube.slang:26-29); ERD
                                    simulations (CAVE_SIM_SHAPE_ERD) wrap every axis with it */
    uint32_t ticks;              /* tick budget (MAX_NUMBER_OF_TICKS_TO_SURVIVE) */
    uint64_t birth;              /* raw bitmask, bit b = count b+1 */
    uint64_t survival;
    uint32_t spawn[3];
    uint32_t spawn_mode; /* CAVE_SPAWN_RANDOM or CAVE_SPAWN_FILLED */
    double density;
    uint32_t seed;
    uint32_t flags; /* CAVE_SIM_* */
} CaveSimSpec;

#define CAVE_SIM_NO_EARLY_EXIT 1u /* run all ticks even after the manager's status 1 */
#define CAVE_SIM_FINAL_HASH 2u    /* compute the FNV-1a canonical hash of the final cells */
/* status 1 when every cell stayed dead during a tick (the superseded test of
 * shaders/source/search/manager.comp:115) instead of "no survivor and no
 * birth" (shaders/slang/search/searchManager.slang:128) */
#define CAVE_SIM_LEGACY_DEATH 4u
/* the elongated rhombic dodecahedron cell shape (shaders/slang/search/
 * searchSimulationERD.slang): 12 face neighbours by y parity + 4 edge
 * neighbours (C ignored), W wraps every axis.  Cube specs ignore W, as the
 * cube search kernel does (searchSimulationCube.slang:26-29). */
#define CAVE_SIM_SHAPE_ERD 8u
/* compute every cell on every tick.  By default the persistent engine
 * computes only the cells within one cell of a cell that has been live so
 * far (all others are 0 and stay 0); the results are the same either way. */
#define CAVE_SIM_FULL_GRID 16u

typedef struct CaveSimResult {
    CaveGridInfo info;       /* GridInfo at completion (manager semantics) */
    uint32_t classification; /* 2 viable, 1 possible, 0 unviable */
    uint32_t ticks_run;
    CaveTickStats last;      /* stats of the last tick run */
    uint64_t final_hash;     /* with CAVE_SIM_FINAL_HASH */
} CaveSimResult;

enum {
    CAVE_ENGINE_AUTO = 0,
    CAVE_ENGINE_PERSISTENT = 1, /* one persistent work-stealing launch for the batch */
    CAVE_ENGINE_PLAIN = 2       /* one step launch per config per tick */
};

typedef struct CaveBatchOptions {
    uint32_t engine;         /* CAVE_ENGINE_* */
    uint32_t block_threads;  /* 0 = autotuned; receives the block size used (persistent engine) */
    CaveTickStats *per_tick; /* nullable: n_specs * max(ticks) entries, row per spec; entries at
                                and after a spec's ticks_run are zero */
    uint32_t per_tick_stride; /* entries per spec row in per_tick (>= max ticks) */
    uint32_t max_parts;      /* most blocks one simulation may be split over (0 = auto, 1 = never split) */
    double device_ms;        /* out: device time of the evaluation */
    struct CaveSimTiming *timing; /* nullable out: one per spec (persistent engine) */
} CaveBatchOptions;

/* Where and when one simulation ran on the device (persistent engine). */
typedef struct CaveSimTiming {
    double start_us; /* device clock at the claim of its first part, after the launch's first claim */
    double end_us;   /* device clock when its result was written */
    uint32_t blocks; /* blocks the simulation was split over */
    uint32_t items_per_thread;
    /* cells its ticks computed, summed over ticks: the active regions' cells
     * (all X*Y*Z per tick with CAVE_SIM_FULL_GRID) */
    uint64_t computed_cells;
} CaveSimTiming;

int cave_batch_run(const CaveSimSpec *specs, size_t n_specs, CaveSimResult *results,
                   CaveBatchOptions *opt);

/* The persistent engine's autotuner table measured so far in this process, as
 * a JSON array of {edge_bucket, flags, threads, items_per_thread,
 * cell_updates_per_s}: one calibration launch per (grid-edge bucket [0,32)
 * [32,48) [48,64) [64,96) [96,128) [128,..), neighbourhood, block size, items
 * per thread).  `flags` is F|E|C for cube grids; ERD grids, whose ticks cost
 * differently, are keyed 16 + F|E, plus 4 when they wrap.  Batch lists add 32
 * for simulations computing every cell (CAVE_SIM_FULL_GRID) and 64 when the
 * rate was measured on a sample of the list's own simulations (buckets where
 * eight of them per resident block stay within 2.5e10 cell-updates) rather
 * than on the synthetic all-counts grid.  No reference counterpart. */
int cave_batch_tuning_report(char *buf, size_t cap);

/* ================================================================== */
/* Reference-compatible --mode search chunk (SearchSystem equivalent)  */
/* ================================================================== */
typedef struct CaveSearch CaveSearch;

typedef struct CaveSearchSpec {
    CaveChunkConfiguration chunk; /* chunk.shapeId selects the cell shape: 0 cube, 1 ERD (records
                                     carry shapeAndGridConfiguration 0 either way, as the reference's) */
    uint32_t neighborhood_flags; /* selected F|E|C (+W); every non-empty subset is a slot, and with W
                                    also the subset with W (searchMode.cpp:56-63) */
    uint32_t max_rule_bits;      /* K (0 = raw bitmask space) */
    int32_t min_max_cell_state;  /* --min-cs */
    int32_t max_max_cell_state;  /* --max-cs-range */
    uint32_t max_ticks;          /* --ticks */
    uint32_t sim_flags;          /* CAVE_SIM_* applied to every evaluation (0 = reference semantics) */
    const char *output_folder;   /* nullable: no JSONL files */
    const char *file_prefix;     /* "<prefix>_{viable,possible,unviable}.jsonl" */
} CaveSearchSpec;

/* SearchResult (src/vulkan/systems/searchSystem.h), rule words in bitmask form */
typedef struct CaveSearchResult {
    CaveSimulationParameters permutation;
    uint32_t ticksSurvived;
    uint32_t totalBirths;
    uint32_t totalDeaths;
    uint32_t maxAliveCount;
    uint32_t finalAliveCount;
    uint32_t category; /* 2 viable, 1 possible, 0 unviable */
} CaveSearchResult;

typedef struct CaveSearchProgress {
    uint64_t total;      /* rule evaluations in this chunk */
    uint64_t completed;
    uint64_t viable, possible, unviable;
} CaveSearchProgress;

int cave_search_create(const CaveSearchSpec *spec, CaveSearch **out);
/* Evaluates up to max_evals pending rules (0 = all) in batched launches. */
int cave_search_step(CaveSearch *h, uint64_t max_evals);
/* Evaluates up to max_evals pending rules (0 = all) of several chunks in shared
 * batched launches, chunk after chunk in list order; each result goes to its
 * own chunk's counters, drain queue and JSONL files, as with cave_search_step.
 * The reference runs the chunks its VRAM budget admits side by side on one
 * GPU (src/modes/searchMode.cpp:349-404); here they share launches. */
int cave_search_step_group(CaveSearch *const *chunks, size_t n_chunks, uint64_t max_evals);
/* PermutationMath::EstimateChunkVramBytes (src/common/permutationMath.h:144-153):
 * the device bytes the reference's SearchSystem takes for one chunk of
 * slot_count slots, (3 x 4 B x ceil(X*Y*Z/8) + 128 B) per slot + 16 MiB.  The
 * sweep admits chunks against --gpu-vram-budget-gb with it. */
uint64_t cave_estimate_chunk_vram_bytes(uint32_t X, uint32_t Y, uint32_t Z, uint32_t slot_count);
int cave_search_is_running(const CaveSearch *h);
int cave_search_progress(const CaveSearch *h, CaveSearchProgress *out);
/* Moves up to cap finished results out; *n receives the count. */
int cave_search_drain(CaveSearch *h, CaveSearchResult *out, size_t cap, size_t *n);
/* Flushes JSONL files; writes the reference summary line into msg. */
int cave_search_write_summary(CaveSearch *h, char *msg, size_t cap);
int cave_search_destroy(CaveSearch *h);

/* ---- rule aggregate (RuleAggregator, src/common/ruleAggregator.h:24-135) ----
 * Results grouped by rule (birth and survival bitmasks, max cell state M,
 * neighbourhood flags, shape id): runs per category and the min / max / sum
 * of ticks survived; written as one JSONL line per rule with the reference's
 * keys, in key order (the reference's map is unordered). */
typedef struct CaveAggregate CaveAggregate;
int cave_aggregate_create(CaveAggregate **out);
/* One result, as RuleAggregator::Record: category 2 viable, 1 possible, 0 unviable. */
int cave_aggregate_record(CaveAggregate *a, const CaveSimulationParameters *params, uint32_t ticks_survived,
                          uint32_t category, uint32_t shape_id);
/* n batch results: each spec's rule (shape id 1 for CAVE_SIM_SHAPE_ERD specs),
 * ticks survived and classification.  Seeds of one rule config fold into one record. */
int cave_aggregate_record_batch(CaveAggregate *a, const CaveSimSpec *specs, const CaveSimResult *results, size_t n);
size_t cave_aggregate_rule_count(const CaveAggregate *a);
int cave_aggregate_write_jsonl(const CaveAggregate *a, const char *path);
int cave_aggregate_destroy(CaveAggregate *a);

/* ================================================================== */
/* Continuous-batching request server                                  */
/* ================================================================== */
typedef struct CaveServer CaveServer;

typedef struct CaveServeOptions {
    uint64_t hbm_budget_bytes; /* admission budget for resident cell state (0 = 200 GB); final cells asked
                                  for are staged in pinned host memory (up to 4 GiB at once) */
    uint32_t max_inflight;     /* requests admitted and not yet finished (0 = 16384) */
    uint32_t split_target_us;  /* a request whose one-block run is estimated longer is split (0 = 1000) */
    uint32_t idle_exit_us;     /* the resident kernel closes after this long without a claim (0 = 50000) */
    uint32_t max_parts;        /* most blocks one request is split over (0 = 64; never above a quarter
                                  of the resident blocks or Z/4) */
} CaveServeOptions;

typedef struct CaveServeCompletion {
    uint64_t request_id;
    CaveSimResult result;
    double submit_s; /* host clock, seconds since cave_serve_create */
    double admit_s;  /* its state was reserved and it was published to the device */
    double finish_s; /* the host saw its completion */
    double run_ms;   /* device time from its first claim to its result */
    uint32_t blocks; /* blocks it ran on */
    int32_t status;  /* CAVE_OK, or CAVE_E_STATE when its parts did not meet in time */
} CaveServeCompletion;

/* Continuous batching: a resident persistent kernel claims requests from a
 * ring in host-coherent memory as the host publishes them, so requests that
 * arrive while others run are admitted (HBM budget, in-flight cap) and start
 * on the next free block.  A request's packed state stays in its own HBM
 * region from admission to completion.  One handle per host thread. */
int cave_serve_create(const CaveServeOptions *opt, CaveServer **out);
/* Queues a request (admitted in arrival order); CAVE_E_BUDGET if it can never
 * fit the budget. */
int cave_serve_submit(CaveServer *s, const CaveSimSpec *spec, uint64_t request_id);
/* Admits what fits, then moves finished requests out (non-blocking). */
int cave_serve_poll(CaveServer *s, CaveServeCompletion *out, size_t cap, size_t *n);
/* Blocks until every submitted request finished (completions stay pollable). */
int cave_serve_drain(CaveServer *s);
int cave_serve_stats(const CaveServer *s, uint64_t *admitted, uint64_t *finished,
                     uint64_t *resident_bytes, uint64_t *peak_resident_bytes);
int cave_serve_destroy(CaveServer *s);

#ifdef __cplusplus
}
#endif
#endif
