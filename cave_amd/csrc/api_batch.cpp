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
#include "common.h"
#include "host_util.h"
#include "step_kernels.h"

namespace cave {

namespace {

constexpr uint32_t kLdsBytesPerCU = 160u * 1024u;
constexpr uint32_t kStaticLds = 1024u;  // red[] + scalars of k_batch

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

int cu_count()
{
    static int cus = 0;
    if (!cus) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, 0) != hipSuccess) return 256;
        cus = p.multiProcessorCount;
    }
    return cus;
}

}  // namespace

int validate_sim(const CaveSimSpec &s)
{
    for (int a = 0; a < 3; a++)
        if (s.dims[a] == 0 || s.dims[a] > 4096) return fail(CAVE_E_INVALID, "sim dims must be in 1..4096");
    if ((uint64_t)s.dims[0] * s.dims[1] * s.dims[2] > 0xFFFFFFFFull)
        return fail(CAVE_E_INVALID, "grid larger than 2^32 cells");
    if (s.max_cell_state < 1 || s.max_cell_state > 16) return fail(CAVE_E_INVALID, "max_cell_state must be 1..16");
    if (s.neighborhood_flags & CAVE_WRAP) return fail(CAVE_E_UNSUPPORTED, "wrap (W) is not supported");
    if (s.neighborhood_flags & ~7u) return fail(CAVE_E_INVALID, "bad neighborhood flags");
    if (s.ticks == 0) return fail(CAVE_E_INVALID, "ticks must be >= 1");
    if (s.spawn_mode > CAVE_SPAWN_FILLED) return fail(CAVE_E_INVALID, "spawn_mode must be random or filled");
    return CAVE_OK;
}

DevJob make_job(const CaveSimSpec &s, uint32_t slot)
{
    DevJob j{};
    j.X = s.dims[0];
    j.Y = s.dims[1];
    j.Z = s.dims[2];
    j.flags = s.neighborhood_flags & 7u;
    j.M = s.max_cell_state;
    const DevRule r = make_rule(s.max_cell_state, s.neighborhood_flags, s.birth, s.survival);
    j.born_lut = r.born_lut;
    j.surv_lut = r.surv_lut;
    j.ticks = s.ticks;
    j.sim_flags = s.flags;
    j.result_slot = slot;
    j.init = make_init(s.dims, s.spawn, s.spawn_mode, s.density, s.seed);
    return j;
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

static uint32_t classify(const CaveGridInfo &info, uint32_t M)
{   // searchSystem.cpp:695-711
    if (info.status == 2u) return 2u;
    return info.numberOfSimulationTicksSurvived > M + 1u ? 1u : 0u;
}

int batch_persistent(const CaveSimSpec *specs, size_t n, CaveSimResult *results, CaveBatchOptions *opt,
                     hipStream_t st)
{
    uint32_t threads = opt && opt->block_threads ? opt->block_threads : 1024u;
    threads = std::max(64u, std::min(1024u, (threads + 63u) / 64u * 64u));
    const uint32_t lds_budget = kLdsBytesPerCU - kStaticLds;

    std::vector<DevJob> jobs(n);
    uint64_t lds_words = 0, arena_words = 0;
    std::vector<uint64_t> final_offsets;
    uint64_t final_words = 0;
    for (size_t i = 0; i < n; i++) {
        jobs[i] = make_job(specs[i], (uint32_t)i);
        jobs[i].zchunks = job_zchunks(jobs[i].X, jobs[i].Y, jobs[i].Z, threads);
        const uint64_t w = job_words(jobs[i].X, jobs[i].Y, jobs[i].Z);
        if (w * 4u <= lds_budget) {
            jobs[i].tier = 0;
            lds_words = std::max(lds_words, w);
        } else {
            jobs[i].tier = 1;
            arena_words = std::max(arena_words, w);
        }
        if (specs[i].flags & CAVE_SIM_FINAL_HASH) {
            final_offsets.push_back(final_words);
            jobs[i].final_slot = (uint32_t)final_offsets.size();
            final_words += cave_cell_words(jobs[i].X, jobs[i].Y, jobs[i].Z);
        }
    }
    // longest jobs first: the shared queue then balances the tail
    std::stable_sort(jobs.begin(), jobs.end(), [](const DevJob &a, const DevJob &b) {
        return (uint64_t)a.X * a.Y * a.Z * a.ticks > (uint64_t)b.X * b.Y * b.Z * b.ticks;
    });

    const uint32_t lds_bytes = (uint32_t)(lds_words * 4u);
    uint32_t per_cu = std::max(1u, 2048u / threads);
    if (lds_bytes) per_cu = std::min(per_cu, std::max(1u, kLdsBytesPerCU / (lds_bytes + kStaticLds)));
    uint32_t blocks = (uint32_t)std::min<uint64_t>((uint64_t)cu_count() * per_cu, n);
    blocks = std::max(1u, blocks);

    DevBuf djobs, dnext, dres, dpt, darena, dfinal, dfoff;
    int rc;
    if ((rc = dev_alloc(djobs, sizeof(DevJob) * n))) return rc;
    if ((rc = dev_alloc(dnext, sizeof(uint32_t)))) return rc;
    if ((rc = dev_alloc(dres, sizeof(DevResult) * n))) return rc;
    const uint32_t stride = opt ? opt->per_tick_stride : 0u;
    const bool want_pt = opt && opt->per_tick && stride;
    if (want_pt && (rc = dev_alloc(dpt, sizeof(CaveTickStats) * n * stride))) return rc;
    if (arena_words && (rc = dev_alloc(darena, sizeof(uint32_t) * arena_words * blocks))) return rc;
    if (final_words) {
        if ((rc = dev_alloc(dfinal, sizeof(uint32_t) * final_words))) return rc;
        if ((rc = dev_alloc(dfoff, sizeof(uint64_t) * final_offsets.size()))) return rc;
        CAVE_HIP(hipMemcpyAsync(dfoff.p, final_offsets.data(), sizeof(uint64_t) * final_offsets.size(),
                                hipMemcpyHostToDevice, st));
    }
    CAVE_HIP(hipMemcpyAsync(djobs.p, jobs.data(), sizeof(DevJob) * n, hipMemcpyHostToDevice, st));
    CAVE_HIP(hipMemsetAsync(dnext.p, 0, sizeof(uint32_t), st));

    BatchArgs a{};
    a.jobs = djobs.as<DevJob>();
    a.njobs = (uint32_t)n;
    a.next_job = dnext.as<uint32_t>();
    a.results = dres.as<DevResult>();
    a.per_tick = want_pt ? dpt.as<CaveTickStats>() : nullptr;
    a.per_tick_stride = stride;
    a.arena = darena.as<uint32_t>();
    a.arena_words = arena_words;
    a.final_cells = dfinal.as<uint32_t>();
    a.final_offsets = dfoff.as<uint64_t>();
    a.lds_words = (uint32_t)lds_words;

    hipEvent_t e0, e1;
    CAVE_HIP(hipEventCreate(&e0));
    CAVE_HIP(hipEventCreate(&e1));
    CAVE_HIP(hipEventRecord(e0, st));
    rc = launch_batch(a, blocks, threads, lds_bytes, st);
    if (rc) return rc;
    CAVE_HIP(hipEventRecord(e1, st));
    std::vector<DevResult> hres(n);
    CAVE_HIP(hipMemcpyAsync(hres.data(), dres.p, sizeof(DevResult) * n, hipMemcpyDeviceToHost, st));
    if (want_pt)
        CAVE_HIP(hipMemcpyAsync(opt->per_tick, dpt.p, sizeof(CaveTickStats) * n * stride, hipMemcpyDeviceToHost, st));
    std::vector<uint32_t> hfinal(final_words);
    if (final_words)
        CAVE_HIP(hipMemcpyAsync(hfinal.data(), dfinal.p, sizeof(uint32_t) * final_words, hipMemcpyDeviceToHost, st));
    CAVE_HIP(hipStreamSynchronize(st));
    float ms = 0.f;
    CAVE_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (opt) opt->device_ms = ms;

    for (size_t i = 0; i < n; i++) {
        CaveSimResult &r = results[i];
        std::memset(&r, 0, sizeof r);
        r.info = hres[i].info;
        r.classification = hres[i].classification;
        r.ticks_run = hres[i].ticks_run;
        r.last = hres[i].last;
    }
    for (const DevJob &j : jobs)
        if (j.final_slot)
            results[j.result_slot].final_hash =
                fnv_ref_layout(hfinal.data() + final_offsets[j.final_slot - 1], j.X, j.Y, j.Z);
    return CAVE_OK;
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
    if (want_pt && (rc = dev_alloc(dpt, sizeof(CaveTickStats) * n * stride))) return rc;
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
        uint32_t *bufs[2] = {b0.as<uint32_t>(), b1.as<uint32_t>()};
        if ((rc = launch_init(bufs[0], g, init, rule.M, st))) return rc;
        PlainState *state = dstate.as<PlainState>() + i;
        const uint32_t *done = &state->done;
        CaveTickStats *pt = want_pt ? dpt.as<CaveTickStats>() + i * stride : nullptr;
        for (uint32_t t = 0; t < s.ticks; t++) {
            uint32_t *in = bufs[t & 1u], *out = bufs[(t + 1u) & 1u];
            rc = use_stream[i] ? launch_stream(in, out, g, rule, plans[i], dparts.as<StatPart>(), st, done)
                               : launch_gather(in, out, g, rule, dparts.as<StatPart>(), st, done);
            if (rc) return rc;
            const uint32_t nparts = use_stream[i] ? plans[i].blocks : gather_blocks(g);
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

}  // namespace cave

using namespace cave;

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
    int rc = engine == CAVE_ENGINE_PLAIN ? batch_plain(specs, n_specs, results, opt, st)
                                         : batch_persistent(specs, n_specs, results, opt, st);
    (void)hipStreamSynchronize(st);
    (void)hipStreamDestroy(st);
    return rc;
}
