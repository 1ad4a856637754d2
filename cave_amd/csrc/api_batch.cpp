// api_batch.cpp -- cave_batch_run: many independent simulations in one call.
//
// ENGINE_PERSISTENT: one launch of the persistent work-stealing kernel
// (batch.hip) for the whole list.  ENGINE_PLAIN: the per-config form of the
// reference's search loop (searchSystem.cpp:543-585: one simulation dispatch
// and one manager dispatch per config per tick), kept as the comparison
// point: it drives the single-grid step kernel config by config, tick by
// tick, with the manager on the device and no host round trip per tick.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <numeric>
#include <vector>

#include "batch.h"
#include "batch_plan.h"
#include "common.h"
#include "host_util.h"
#include "step_kernels.h"

namespace cave {

namespace {

struct DevBuf {
    void *p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    template <typename T>
    T *as() const { return static_cast<T *>(p); }
};

int dev_alloc(DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) {
        b.p = nullptr;
        return fail(e == hipErrorOutOfMemory ? CAVE_E_NOMEM : CAVE_E_DEVICE,
                    std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    return CAVE_OK;
}

}  // namespace

int validate_sim(const CaveSimSpec &s)
{
    for (int a = 0; a < 3; a++)
        if (s.dims[a] == 0 || s.dims[a] > 4096) return fail(CAVE_E_INVALID, "sim dims must be in 1..4096");
    if ((uint64_t)s.dims[0] * s.dims[1] * s.dims[2] > 0xFFFFFFFFull)
        return fail(CAVE_E_INVALID, "grid larger than 2^32 cells");
    if (s.max_cell_state < 1 || s.max_cell_state > 16) return fail(CAVE_E_INVALID, "max_cell_state must be 1..16");
    // W: the cube search kernel ignores it (searchSimulationCube.slang:26-29), the ERD one wraps
    if (s.neighborhood_flags & ~15u) return fail(CAVE_E_INVALID, "bad neighborhood flags");
    if (s.flags & ~(CAVE_SIM_NO_EARLY_EXIT | CAVE_SIM_FINAL_HASH | CAVE_SIM_LEGACY_DEATH | CAVE_SIM_SHAPE_ERD |
                    CAVE_SIM_FULL_GRID))
        return fail(CAVE_E_INVALID, "unknown sim flags");
    if (s.ticks == 0) return fail(CAVE_E_INVALID, "ticks must be >= 1");
    if (s.spawn_mode == CAVE_SPAWN_BUFFER || s.spawn_mode > CAVE_SPAWN_RENDER_RANDOM)
        return fail(CAVE_E_INVALID, "spawn_mode must be random, filled or render-random");
    return CAVE_OK;
}

uint64_t fnv_ref_layout(const uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z)
{   // FNV-1a-64 over the canonical unpadded packing (x + X(y + Yz), two cells per byte)
    const uint32_t upr = (X + 7u) / 8u;
    uint64_t h = 1469598103934665603ull;
    const uint64_t n = (uint64_t)X * Y * Z;
    uint64_t i = 0;
    uint32_t byte = 0;
    for (uint32_t z = 0; z < Z; z++)
        for (uint32_t y = 0; y < Y; y++)
            for (uint32_t x = 0; x < X; x++, i++) {
                const uint32_t v = (cells[x / 8u + (size_t)upr * (y + (size_t)Y * z)] >> (4u * (x % 8u))) & 15u;
                if ((i & 1u) == 0) {
                    byte = v;
                    if (i + 1 == n) {
                        h ^= byte;
                        h *= 1099511628211ull;
                    }
                } else {
                    byte |= v << 4;
                    h ^= byte;
                    h *= 1099511628211ull;
                }
            }
    return h;
}

// FNV-1a-64 of the canonical byte stream, which is the ERD linear layout itself
uint64_t fnv_linear(const uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z)
{
    const uint64_t n = (uint64_t)X * Y * Z;
    uint64_t h = 1469598103934665603ull;
    for (uint64_t i = 0; i < n; i += 2) {
        uint32_t byte = (cells[i / 8u] >> (4u * (i % 8u))) & 15u;
        if (i + 1 < n) byte |= ((cells[(i + 1) / 8u] >> (4u * ((i + 1) % 8u))) & 15u) << 4;
        h ^= byte;
        h *= 1099511628211ull;
    }
    return h;
}

static uint32_t classify(const CaveGridInfo &info, uint32_t M)
{   // searchSystem.cpp:695-711
    if (info.status == 2u) return 2u;
    return info.numberOfSimulationTicksSurvived > M + 1u ? 1u : 0u;
}

int batch_persistent(const CaveSimSpec *specs, size_t n, CaveSimResult *results, CaveBatchOptions *opt,
                     hipStream_t st)
{
    BatchKnobs k;
    k.threads = opt ? opt->block_threads : 0u;
    k.max_parts = opt ? opt->max_parts : 0u;
    BatchPlan plan;
    int rc = plan_batch(specs, n, k, plan);
    if (rc) return rc;
    if (opt) opt->block_threads = plan.threads;
    return run_plan(plan, n, results, opt, st);
}

int batch_plain(const CaveSimSpec *specs, size_t n, CaveSimResult *results, CaveBatchOptions *opt, hipStream_t st)
{
    // buffers sized for the largest config
    size_t max_words = 0;
    uint32_t max_parts = 1;
    std::vector<GridGeom> geo(n);
    std::vector<int> use_stream(n);
    std::vector<StreamPlan> plans(n);
    for (size_t i = 0; i < n; i++) {
        GridGeom g{specs[i].dims[0], specs[i].dims[1], specs[i].dims[2], 4u * ((specs[i].dims[0] + 31u) / 32u), 1u};
        if (specs[i].flags & CAVE_SIM_SHAPE_ERD) {  // linear layout, ERD kernel
            g.planar = 0;
            geo[i] = g;
            use_stream[i] = 0;
            max_parts = std::max(max_parts, erd_blocks(g.X, g.Y, g.Z));
            max_words = std::max(max_words, cave_erd_cell_words(g.X, g.Y, g.Z));
            continue;
        }
        use_stream[i] = stream_supported(g);
        if (use_stream[i]) {
            plans[i] = stream_plan(g, 0, 0, 0);
            max_parts = std::max(max_parts, plans[i].blocks);
        } else {
            g.pitch = ref_words_per_row(g.X);
            g.planar = 0;
            max_parts = std::max(max_parts, gather_blocks(g));
        }
        geo[i] = g;
        max_words = std::max(max_words, (size_t)g.pitch * g.Y * g.Z);
    }
    const uint32_t stride = opt ? opt->per_tick_stride : 0u;
    const bool want_pt = opt && opt->per_tick && stride;
    DevBuf b0, b1, dparts, dstate, dpt;
    int rc;
    if ((rc = dev_alloc(b0, max_words * 4))) return rc;
    if ((rc = dev_alloc(b1, max_words * 4))) return rc;
    if ((rc = dev_alloc(dparts, sizeof(StatPart) * max_parts))) return rc;
    if ((rc = dev_alloc(dstate, sizeof(PlainState) * n))) return rc;
    if (want_pt) {  // rows past a simulation's ticks_run come back zero
        if ((rc = dev_alloc(dpt, sizeof(CaveTickStats) * n * stride))) return rc;
        CAVE_HIP(hipMemsetAsync(dpt.p, 0, sizeof(CaveTickStats) * n * stride, st));
    }
    CAVE_HIP(hipMemsetAsync(dstate.p, 0, sizeof(PlainState) * n, st));

    hipEvent_t e0, e1;
    CAVE_HIP(hipEventCreate(&e0));
    CAVE_HIP(hipEventCreate(&e1));
    CAVE_HIP(hipEventRecord(e0, st));
    std::vector<uint64_t> hashes(n, 0);
    std::vector<PlainState> hs(n);
    for (size_t i = 0; i < n; i++) {
        const CaveSimSpec &s = specs[i];
        const GridGeom &g = geo[i];
        const DevRule rule = make_rule(s.max_cell_state, s.neighborhood_flags, s.birth, s.survival);
        const DevInit init = make_init(s.dims, s.spawn, s.spawn_mode, s.density, s.seed);
        const bool erd = (s.flags & CAVE_SIM_SHAPE_ERD) != 0;
        const bool wrap = erd && (s.neighborhood_flags & CAVE_WRAP);
        uint32_t *bufs[2] = {b0.as<uint32_t>(), b1.as<uint32_t>()};
        if ((rc = erd ? launch_init_erd(bufs[0], g.X, g.Y, g.Z, init, rule.M, st)
                      : launch_init(bufs[0], g, init, rule.M, st)))
            return rc;
        PlainState *state = dstate.as<PlainState>() + i;
        const uint32_t *done = &state->done;
        CaveTickStats *pt = want_pt ? dpt.as<CaveTickStats>() + i * stride : nullptr;
        for (uint32_t t = 0; t < s.ticks; t++) {
            uint32_t *in = bufs[t & 1u], *out = bufs[(t + 1u) & 1u];
            rc = erd            ? launch_erd(in, out, g.X, g.Y, g.Z, rule, wrap, dparts.as<StatPart>(), st, done)
                 : use_stream[i] ? launch_stream(in, out, g, rule, plans[i], dparts.as<StatPart>(), st, done)
                                 : launch_gather(in, out, g, rule, dparts.as<StatPart>(), st, done);
            if (rc) return rc;
            const uint32_t nparts = erd ? erd_blocks(g.X, g.Y, g.Z) : use_stream[i] ? plans[i].blocks : gather_blocks(g);
            if ((rc = launch_plain_manager(dparts.as<StatPart>(), nparts, state, s.ticks, s.flags, pt,
                                           g.X * g.Y * g.Z, st)))
                return rc;
        }
        if (s.flags & CAVE_SIM_FINAL_HASH) {
            CAVE_HIP(hipMemcpyAsync(&hs[i], state, sizeof(PlainState), hipMemcpyDeviceToHost, st));
            CAVE_HIP(hipStreamSynchronize(st));
            const uint32_t upr = ref_words_per_row(g.X);
            std::vector<uint32_t> cells((size_t)upr * g.Y * g.Z);
            uint32_t *fin = bufs[hs[i].info.numberOfSimulationTicksSurvived & 1u];
            uint32_t *idle = bufs[(hs[i].info.numberOfSimulationTicksSurvived + 1u) & 1u];
            if (erd) {
                std::vector<uint32_t> lin(cave_erd_cell_words(g.X, g.Y, g.Z));
                CAVE_HIP(hipMemcpyAsync(lin.data(), fin, lin.size() * 4, hipMemcpyDeviceToHost, st));
                CAVE_HIP(hipStreamSynchronize(st));
                hashes[i] = fnv_linear(lin.data(), g.X, g.Y, g.Z);
                continue;
            }
            if (g.planar) {
                if ((rc = launch_planar_to_ref(fin, idle, g, st))) return rc;
                CAVE_HIP(hipMemcpyAsync(cells.data(), idle, cells.size() * 4, hipMemcpyDeviceToHost, st));
            } else {
                CAVE_HIP(hipMemcpy2DAsync(cells.data(), (size_t)upr * 4, fin, (size_t)g.pitch * 4, (size_t)upr * 4,
                                          (size_t)g.Y * g.Z, hipMemcpyDeviceToHost, st));
            }
            CAVE_HIP(hipStreamSynchronize(st));
            hashes[i] = fnv_ref_layout(cells.data(), g.X, g.Y, g.Z);
        }
    }
    CAVE_HIP(hipEventRecord(e1, st));
    CAVE_HIP(hipMemcpyAsync(hs.data(), dstate.p, sizeof(PlainState) * n, hipMemcpyDeviceToHost, st));
    if (want_pt)
        CAVE_HIP(hipMemcpyAsync(opt->per_tick, dpt.p, sizeof(CaveTickStats) * n * stride, hipMemcpyDeviceToHost, st));
    CAVE_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    CAVE_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (opt) opt->device_ms = ms;
    for (size_t i = 0; i < n; i++) {
        CaveSimResult &r = results[i];
        std::memset(&r, 0, sizeof r);
        r.info = hs[i].info;
        r.classification = classify(hs[i].info, specs[i].max_cell_state);
        r.ticks_run = hs[i].info.numberOfSimulationTicksSurvived;
        r.last = hs[i].last;
        r.final_hash = hashes[i];
    }
    return CAVE_OK;
}

std::string tuning_report();

// A batch runs as one persistent launch per cell shape among its specs
// (kBatchShape*: cube, ERD, wrapping ERD -- the kernel is instantiated per
// shape), unless the caller asked for the per-config engine.
// Runs each group of the batch on its engine; results, per-tick rows and
// timing scattered back to spec order.  The device time is the groups' sum.
int batch_grouped(const CaveSimSpec *specs, size_t n, CaveSimResult *results, CaveBatchOptions *opt, hipStream_t st)
{
    std::vector<size_t> idx[3];
    for (size_t i = 0; i < n; i++) idx[spec_shape(specs[i])].push_back(i);
    for (int g = 0; g < 3; g++)
        if (idx[g].size() == n) return batch_persistent(specs, n, results, opt, st);  // one group: no scatter
    const uint32_t stride = opt ? opt->per_tick_stride : 0u;
    const bool want_pt = opt && opt->per_tick && stride;
    double total_ms = 0;
    for (int g = 0; g < 3; g++) {
        const std::vector<size_t> &ix = idx[g];
        if (ix.empty()) continue;
        std::vector<CaveSimSpec> sub(ix.size());
        for (size_t k = 0; k < ix.size(); k++) sub[k] = specs[ix[k]];
        std::vector<CaveSimResult> res(ix.size());
        std::vector<CaveTickStats> pt(want_pt ? ix.size() * stride : 0);
        std::vector<CaveSimTiming> tm(opt && opt->timing ? ix.size() : 0);
        CaveBatchOptions o{};
        if (opt) o = *opt;
        o.per_tick = want_pt ? pt.data() : nullptr;
        o.timing = tm.empty() ? nullptr : tm.data();
        int rc = batch_persistent(sub.data(), sub.size(), res.data(), &o, st);
        if (rc) return rc;
        total_ms += o.device_ms;
        if (opt) opt->block_threads = o.block_threads;
        for (size_t k = 0; k < ix.size(); k++) {
            results[ix[k]] = res[k];
            if (want_pt) std::copy(pt.begin() + k * stride, pt.begin() + (k + 1) * stride, opt->per_tick + ix[k] * stride);
            if (!tm.empty()) opt->timing[ix[k]] = tm[k];
        }
    }
    if (opt) opt->device_ms = total_ms;
    return CAVE_OK;
}

}  // namespace cave

using namespace cave;

extern "C" int cave_batch_tuning_report(char *buf, size_t cap)
{
    if (!buf || cap == 0) return fail(CAVE_E_INVALID, "null buffer");
    const std::string s = tuning_report();
    if (s.size() + 1 > cap) return fail(CAVE_E_INVALID, "buffer too small for the tuning report");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return CAVE_OK;
}

extern "C" int cave_batch_run(const CaveSimSpec *specs, size_t n_specs, CaveSimResult *results,
                              CaveBatchOptions *opt)
{
    if (!specs || !results) return fail(CAVE_E_INVALID, "null argument");
    if (n_specs == 0) return CAVE_OK;
    if (n_specs > 0xFFFFFFFFull) return fail(CAVE_E_INVALID, "too many specs");
    for (size_t i = 0; i < n_specs; i++) {
        int rc = validate_sim(specs[i]);
        if (rc) return fail(rc, "spec " + std::to_string(i) + ": " + cave_last_error());
    }
    if (opt && opt->per_tick && opt->per_tick_stride) {
        for (size_t i = 0; i < n_specs; i++)
            if (specs[i].ticks > opt->per_tick_stride) return fail(CAVE_E_INVALID, "per_tick_stride < ticks");
    }
    const uint32_t engine = opt ? opt->engine : CAVE_ENGINE_AUTO;
    if (engine > CAVE_ENGINE_PLAIN) return fail(CAVE_E_INVALID, "bad engine");
    hipStream_t st;
    CAVE_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    const int rc = engine == CAVE_ENGINE_PLAIN ? batch_plain(specs, n_specs, results, opt, st)
                                               : batch_grouped(specs, n_specs, results, opt, st);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    return rc;
}
