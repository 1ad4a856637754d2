// api_step.cpp -- cave_step_*: the single-grid (headless) step path.
//
// Replaces the headless step loop (src/modes/renderingMode.cpp:166-343 driving
// ComputeSystem::Tick, src/vulkan/systems/computeSystem.cpp:144-395) with a
// HIP stream owning two resident cell buffers (pointer-swap ping-pong, no
// copy-back as in searchSystem.cpp:599-616).  The transition is the search
// kernel's (searchSimulationCube.slang:78-247), which the render path's
// (rayMarchCompute.slang:150-178) equals on non-wrapping grids; with the W
// flag every axis wraps (rayMarchCompute.slang:80-108,
// simulationRenderer.cpp:595) on the planar streaming kernel (grids under
// 1 GiB per buffer) or the gather kernel.  With
// CAVE_STEP_GRAPHS=1 a run of ticks is one HIP graph of step launches
// (captured once per chunk length and buffer parity, replayed after).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "common.h"
#include "host_util.h"
#include "step_kernels.h"

using namespace cave;

struct CaveStep {
    GridGeom g{};
    DevRule rule{};
    CaveSimulationParameters params{};
    int kernel = CAVE_KERNEL_GATHER;
    bool wrap = false;
    uint32_t shape = 0;  // 0 cube, 1 elongated rhombic dodecahedron (linear layout)
    StreamPlan plan{};
    uint32_t nparts = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint32_t *cur = nullptr, *nxt = nullptr;
    StatPart *partials = nullptr;   // [chunk ticks][nparts]
    // [chunk ticks] of per-tick rows in pinned, host-coherent memory that the
    // reduction kernel writes directly; it then sets *hflag (host-coherent),
    // which the host polls instead of waiting for the stream
    CaveTickStats *hstats = nullptr, *hstats_dev = nullptr;
    uint32_t *hflag = nullptr, *hflag_dev = nullptr;
    uint32_t *dcount = nullptr;  // device counter of the reduction's finished blocks
    uint32_t chunk = 0;
    // instantiated tick loops: one graph per (ticks, which buffer is current)
    struct TickGraph {
        uint32_t ticks;
        const uint32_t *cur;
        hipGraphExec_t exec;
    };
    std::vector<TickGraph> graphs;
    bool use_graphs = false;
};

namespace {

constexpr uint32_t kChunkTicks = 1024;
// Device bytes of block partials one chunk of ticks may hold: a chunk is
// min(kChunkTicks, kPartialBudget / (blocks per tick * sizeof(StatPart))) ticks.
constexpr size_t kPartialBudget = 64u << 20;

uint32_t chunk_ticks(uint32_t nparts)
{
    const size_t per_tick = (size_t)std::max<uint32_t>(nparts, 1u) * sizeof(StatPart);
    return (uint32_t)std::clamp<size_t>(kPartialBudget / per_tick, 1, kChunkTicks);
}

int parse_rule(const CaveSimulationParameters &p, DevRule *out, bool *wrap, uint32_t *shape)
{
    const uint32_t flags = uint32_t(p.survivalAndNeighborhoodRules >> 60);
    *wrap = (flags & CAVE_WRAP) != 0;
    *shape = uint32_t(p.shapeAndGridConfiguration & 0x1Fu);  // searchSystem.cpp:198-199
    if (*shape > 1u) return fail(CAVE_E_UNSUPPORTED, "shape ids are 0 (cube) and 1 (ERD)");
    const uint32_t M = uint32_t(p.birthAndMaxCellStateRules >> 60) + 1u;
    *out = make_rule(M, flags, p.birthAndMaxCellStateRules & kPermMask, p.survivalAndNeighborhoodRules & kPermMask);
    return CAVE_OK;
}

void drop_graphs(CaveStep *h)
{
    for (auto &g : h->graphs) (void)hipGraphExecDestroy(g.exec);
    h->graphs.clear();
}

// Launches the n step kernels of a chunk (ticks t write partials row t).
int launch_ticks(CaveStep *h, uint32_t n, uint32_t *cur, uint32_t *nxt)
{
    for (uint32_t t = 0; t < n; t++) {
        StatPart *parts = h->partials + (size_t)t * h->nparts;
        int rc = h->kernel == CAVE_KERNEL_ERD
                     ? launch_erd(cur, nxt, h->g.X, h->g.Y, h->g.Z, h->rule, h->wrap, parts, h->stream)
                 : h->kernel != CAVE_KERNEL_GATHER
                     ? launch_stream(cur, nxt, h->g, h->rule, h->plan, parts, h->stream, nullptr, h->wrap)
                     : launch_gather(cur, nxt, h->g, h->rule, parts, h->stream, nullptr, h->wrap);
        if (rc) return rc;
        std::swap(cur, nxt);
    }
    return CAVE_OK;
}

// The chunk's launches as one instantiated graph, captured on first use.
int tick_graph(CaveStep *h, uint32_t n, hipGraphExec_t *exec)
{
    for (const auto &g : h->graphs)
        if (g.ticks == n && g.cur == h->cur) {
            *exec = g.exec;
            return CAVE_OK;
        }
    if (h->graphs.size() >= 8) drop_graphs(h);
    hipGraph_t graph = nullptr;
    CAVE_HIP(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = launch_ticks(h, n, h->cur, h->nxt);
    const hipError_t e = hipStreamEndCapture(h->stream, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return fail(CAVE_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t x = nullptr;
    const hipError_t ie = hipGraphInstantiate(&x, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) return fail(CAVE_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    h->graphs.push_back({n, h->cur, x});
    *exec = x;
    return CAVE_OK;
}

void release(CaveStep *h)
{
    if (!h) return;
    drop_graphs(h);
    if (h->cur) (void)hipFree(h->cur);
    if (h->nxt) (void)hipFree(h->nxt);
    if (h->partials) (void)hipFree(h->partials);
    if (h->dcount) (void)hipFree(h->dcount);
    if (h->hstats) (void)hipHostFree(h->hstats);
    if (h->hflag) (void)hipHostFree(h->hflag);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

size_t dev_words(const CaveStep *h)
{
    if (h->shape == 1u) return cave_erd_cell_words(h->g.X, h->g.Y, h->g.Z);
    return (size_t)h->g.pitch * h->g.Y * h->g.Z;
}

}  // namespace

extern "C" {

int cave_step_create(const CaveStepSpec *spec, CaveStep **out)
{
    if (!spec || !out) return fail(CAVE_E_INVALID, "null argument");
    *out = nullptr;
    for (int a = 0; a < 3; a++)
        if (spec->dims[a] == 0 || spec->dims[a] > 65536)
            return fail(CAVE_E_INVALID, "grid dimensions must be in 1..65536");
    const uint64_t cells = (uint64_t)spec->dims[0] * spec->dims[1] * spec->dims[2];
    if (spec->spawn_mode > CAVE_SPAWN_RENDER_RANDOM) return fail(CAVE_E_INVALID, "bad spawn_mode");
    if (spec->spawn_mode == CAVE_SPAWN_BUFFER && !spec->initial_cells)
        return fail(CAVE_E_INVALID, "CAVE_SPAWN_BUFFER needs initial_cells");
    if (spec->kernel > CAVE_KERNEL_ERD) return fail(CAVE_E_INVALID, "bad kernel selector");

    CaveStep *h = new (std::nothrow) CaveStep();
    if (!h) return fail(CAVE_E_NOMEM, "host allocation failed");
    int rc = parse_rule(spec->params, &h->rule, &h->wrap, &h->shape);
    if (rc) { release(h); return rc; }
    if (h->shape == 1u && cells > 0xFFFFFFFFull) {  // the ERD kernel indexes cells with 32 bits
        release(h);
        return fail(CAVE_E_INVALID, "ERD grids are limited to 2^32 - 1 cells");
    }
    if ((h->shape == 1u) != (spec->kernel == CAVE_KERNEL_ERD) && spec->kernel != CAVE_KERNEL_AUTO) {
        release(h);
        return fail(CAVE_E_INVALID, "CAVE_KERNEL_ERD runs exactly the ERD shape (id 1)");
    }
    h->params = spec->params;
    h->g.X = spec->dims[0];
    h->g.Y = spec->dims[1];
    h->g.Z = spec->dims[2];

    GridGeom sg = h->g;
    sg.pitch = 4u * ((h->g.X + 31u) / 32u);
    // W: the planar streaming kernel wraps grids of under 1 GiB per buffer;
    // the nibble layout and larger grids wrap on the gather kernel
    GridGeom wg = sg;
    wg.planar = 1u;
    const bool wrap_ok = stream_wrap_supported(wg) && spec->kernel != CAVE_KERNEL_STREAM_NIBBLE;
    const bool stream_ok = stream_supported(sg) && h->shape == 0u && (!h->wrap || wrap_ok);
    const bool want_stream = spec->kernel == CAVE_KERNEL_STREAM || spec->kernel == CAVE_KERNEL_STREAM_NIBBLE;
    if (want_stream && !stream_ok) {
        release(h);
        return fail(CAVE_E_UNSUPPORTED, h->wrap ? "wrapping grids run on the planar streaming kernel (under 1 GiB "
                                                  "per buffer) or the gather kernel"
                                                : "the streaming kernel needs X <= 2048");
    }
    if (h->shape == 1u) h->kernel = CAVE_KERNEL_ERD;
    else if (spec->kernel == CAVE_KERNEL_GATHER || !stream_ok) h->kernel = CAVE_KERNEL_GATHER;
    else h->kernel = spec->kernel == CAVE_KERNEL_STREAM_NIBBLE ? CAVE_KERNEL_STREAM_NIBBLE : CAVE_KERNEL_STREAM;
    if (h->kernel == CAVE_KERNEL_ERD) {
        h->g.pitch = 0;
        h->nparts = erd_blocks(h->g.X, h->g.Y, h->g.Z);
    } else if (h->kernel != CAVE_KERNEL_GATHER) {
        sg.planar = h->kernel == CAVE_KERNEL_STREAM ? 1u : 0u;
        h->g = sg;
        // CAVE_STREAM_PLAN="rows_per_lane,z_planes,waves" overrides the default tiling
        uint32_t rpl = 0, tz = 0, waves = 0;
        if (const char *env = std::getenv("CAVE_STREAM_PLAN")) std::sscanf(env, "%u,%u,%u", &rpl, &tz, &waves);
        h->plan = stream_plan(h->g, rpl, tz, waves);
        h->nparts = h->plan.blocks;
    } else {
        h->g.pitch = ref_words_per_row(h->g.X);
        h->nparts = gather_blocks(h->g);
    }

    const size_t bytes = dev_words(h) * sizeof(uint32_t);
#define CK(expr)                                                                 \
    do {                                                                         \
        hipError_t _e = (expr);                                                  \
        if (_e != hipSuccess) {                                                  \
            release(h);                                                          \
            return fail(_e == hipErrorOutOfMemory ? CAVE_E_NOMEM : CAVE_E_DEVICE, \
                        std::string(#expr) + ": " + hipGetErrorString(_e));     \
        }                                                                        \
    } while (0)
    CK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    CK(hipEventCreate(&h->ev0));
    CK(hipEventCreate(&h->ev1));
    CK(hipMalloc(&h->cur, bytes));
    CK(hipMalloc(&h->nxt, bytes));
    CK(hipMemsetAsync(h->nxt, 0, bytes, h->stream));
    h->chunk = chunk_ticks(h->nparts);
    // CAVE_STEP_GRAPHS=1 replays each run of ticks as a captured graph.  Off by
    // default: the eager launches already run back to back on the device (the
    // same device time per tick at 49^3 .. 512^3), and a capture costs host
    // time the first time a run length is seen.
    if (const char *env = std::getenv("CAVE_STEP_GRAPHS")) h->use_graphs = env[0] == '1';
    CK(hipMalloc(&h->partials, sizeof(StatPart) * (size_t)h->nparts * h->chunk));
    // Per-tick rows go device -> pinned staging -> caller: the reduction
    // kernel writes them straight into host-coherent memory and raises a
    // flag there, so a call ends without a device-to-host copy and without
    // the stream synchronisation's wake-up latency (a copy straight into a
    // caller's pageable buffer of this size costs the runtime milliseconds).
    CK(hipMalloc(&h->dcount, sizeof(uint32_t)));
    CK(hipMemsetAsync(h->dcount, 0, sizeof(uint32_t), h->stream));
    CK(hipHostMalloc((void **)&h->hstats, sizeof(CaveTickStats) * h->chunk,
                     hipHostMallocCoherent | hipHostMallocMapped));
    CK(hipHostMalloc((void **)&h->hflag, sizeof(uint32_t), hipHostMallocCoherent | hipHostMallocMapped));
    CK(hipHostGetDevicePointer((void **)&h->hstats_dev, h->hstats, 0));
    CK(hipHostGetDevicePointer((void **)&h->hflag_dev, h->hflag, 0));
    if (spec->spawn_mode == CAVE_SPAWN_BUFFER) {
        CK(hipStreamSynchronize(h->stream));
        rc = cave_step_upload(h, spec->initial_cells, dev_words(h));
        if (rc) { release(h); return rc; }
    } else {
        const uint32_t dims[3] = {h->g.X, h->g.Y, h->g.Z};
        const DevInit init = make_init(dims, spec->spawn, spec->spawn_mode, spec->density, spec->seed);
        rc = h->shape == 1u ? launch_init_erd(h->cur, h->g.X, h->g.Y, h->g.Z, init, h->rule.M, h->stream)
                            : launch_init(h->cur, h->g, init, h->rule.M, h->stream);
        if (rc) { release(h); return rc; }
    }
    CK(hipStreamSynchronize(h->stream));
#undef CK
    *out = h;
    return CAVE_OK;
}

int cave_step_run(CaveStep *h, uint32_t ticks, CaveTickStats *per_tick, double *device_ms)
{
    if (!h) return fail(CAVE_E_INVALID, "null handle");
    double total_ms = 0.0;
    uint32_t done = 0;
    while (done < ticks) {
        const uint32_t n = std::min(ticks - done, h->chunk);
        hipGraphExec_t exec = nullptr;
        if (h->use_graphs) {
            int rc = tick_graph(h, n, &exec);
            if (rc) return rc;
        }
        CAVE_HIP(hipEventRecord(h->ev0, h->stream));
        if (exec) {
            CAVE_HIP(hipGraphLaunch(exec, h->stream));
        } else {
            int rc = launch_ticks(h, n, h->cur, h->nxt);
            if (rc) return rc;
        }
        if (n & 1u) std::swap(h->cur, h->nxt);
        CAVE_HIP(hipEventRecord(h->ev1, h->stream));
        if (per_tick) {
            __atomic_store_n(h->hflag, 0u, __ATOMIC_RELEASE);
            int rc = launch_reduce_ticks(h->partials, h->nparts, n, h->hstats_dev, h->stream, h->dcount, h->hflag_dev);
            if (rc) return rc;
            // The rows are complete once the flag is up.  A failed launch
            // never raises it: the stream query reports the error (or the
            // stream's end without the flag) instead.
            for (uint32_t spin = 1;; spin++) {
                if (__atomic_load_n(h->hflag, __ATOMIC_ACQUIRE)) break;
                if ((spin & 1023u) == 0u) {
                    const hipError_t q = hipStreamQuery(h->stream);
                    if (q == hipSuccess) {
                        if (__atomic_load_n(h->hflag, __ATOMIC_ACQUIRE)) break;
                        (void)hipMemset(h->dcount, 0, sizeof(uint32_t));  // for the next call
                        return fail(CAVE_E_DEVICE, "the per-tick reduction ended without its completion flag");
                    }
                    if (q != hipErrorNotReady) CAVE_HIP(q);
                }
            }
            std::memcpy(per_tick + done, h->hstats, sizeof(CaveTickStats) * n);
            CAVE_HIP(hipEventSynchronize(h->ev1));
        } else {
            CAVE_HIP(hipStreamSynchronize(h->stream));
        }
        float ms = 0.f;
        CAVE_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        total_ms += ms;
        done += n;
    }
    if (device_ms) *device_ms = total_ms;
    return CAVE_OK;
}

int cave_step_download(CaveStep *h, uint32_t *cells, size_t n_words)
{
    if (!h || !cells) return fail(CAVE_E_INVALID, "null argument");
    if (h->shape == 1u) {  // the ERD linear layout is stored as is
        if (n_words < dev_words(h)) return fail(CAVE_E_INVALID, "buffer too small");
        CAVE_HIP(hipMemcpyAsync(cells, h->cur, dev_words(h) * 4, hipMemcpyDeviceToHost, h->stream));
        CAVE_HIP(hipStreamSynchronize(h->stream));
        return CAVE_OK;
    }
    const uint32_t upr = ref_words_per_row(h->g.X);
    if (n_words < (size_t)upr * h->g.Y * h->g.Z) return fail(CAVE_E_INVALID, "buffer too small");
    if (h->g.planar) {
        // convert into the idle ping-pong buffer, then copy the reference layout out
        int rc = launch_planar_to_ref(h->cur, h->nxt, h->g, h->stream);
        if (rc) return rc;
        CAVE_HIP(hipMemcpyAsync(cells, h->nxt, (size_t)upr * h->g.Y * h->g.Z * 4, hipMemcpyDeviceToHost, h->stream));
    } else {
        CAVE_HIP(hipMemcpy2DAsync(cells, (size_t)upr * 4, h->cur, (size_t)h->g.pitch * 4, (size_t)upr * 4,
                                  (size_t)h->g.Y * h->g.Z, hipMemcpyDeviceToHost, h->stream));
    }
    CAVE_HIP(hipStreamSynchronize(h->stream));
    return CAVE_OK;
}

int cave_step_upload(CaveStep *h, const uint32_t *cells, size_t n_words)
{
    if (!h || !cells) return fail(CAVE_E_INVALID, "null argument");
    if (h->shape == 1u) {
        if (n_words < dev_words(h)) return fail(CAVE_E_INVALID, "buffer too small");
        CAVE_HIP(hipMemcpyAsync(h->cur, cells, dev_words(h) * 4, hipMemcpyHostToDevice, h->stream));
        CAVE_HIP(hipStreamSynchronize(h->stream));
        return CAVE_OK;
    }
    const uint32_t upr = ref_words_per_row(h->g.X);
    if (n_words < (size_t)upr * h->g.Y * h->g.Z) return fail(CAVE_E_INVALID, "buffer too small");
    if (h->g.planar) {
        CAVE_HIP(hipMemcpyAsync(h->nxt, cells, (size_t)upr * h->g.Y * h->g.Z * 4, hipMemcpyHostToDevice, h->stream));
        int rc = launch_ref_to_planar(h->nxt, h->cur, h->g, h->stream);
        if (rc) return rc;
    } else {
        CAVE_HIP(hipMemsetAsync(h->cur, 0, dev_words(h) * 4, h->stream));
        CAVE_HIP(hipMemcpy2DAsync(h->cur, (size_t)h->g.pitch * 4, cells, (size_t)upr * 4, (size_t)upr * 4,
                                  (size_t)h->g.Y * h->g.Z, hipMemcpyHostToDevice, h->stream));
        int rc = launch_sanitize_nibbles(h->cur, h->g, h->stream);
        if (rc) return rc;
    }
    CAVE_HIP(hipStreamSynchronize(h->stream));
    return CAVE_OK;
}

int cave_step_set_params(CaveStep *h, const CaveSimulationParameters *params)
{
    if (!h || !params) return fail(CAVE_E_INVALID, "null argument");
    DevRule r{};
    bool wrap = false;
    uint32_t shape = 0;
    int rc = parse_rule(*params, &r, &wrap, &shape);
    if (rc) return rc;
    if (wrap != h->wrap || shape != h->shape)
        return fail(CAVE_E_STATE, "the W flag and the shape cannot change on a live handle (they select the kernel)");
    drop_graphs(h);  // the rule is baked into the captured launches
    h->rule = r;
    h->params = *params;
    return CAVE_OK;
}

int cave_step_kernel(const CaveStep *h) { return h ? h->kernel : -1; }

int cave_step_destroy(CaveStep *h)
{
    if (h && h->stream) (void)hipStreamSynchronize(h->stream);
    release(h);
    return CAVE_OK;
}

}  // extern "C"
