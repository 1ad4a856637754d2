// api_static.hip -- cave_step_run_static: ticks of the single-grid step on
// caller-owned device buffers, with no device allocation at all.
//
// cave_step_run keeps its block partials and per-tick statistics in buffers
// the handle allocates.  Callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc) call this entry point instead: the cells and the statistics are
// the caller's buffers, and the only scratch, one tick's block partials,
// lives in the file-scope device array `scratch` below.  Each device holds
// its own copy of that array from the library's first use on it until the
// process ends; every call on a device shares it, so calls that overlap on
// one device (concurrent streams, other host threads) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "common.h"
#include "step_kernels.h"

using namespace cave;

namespace {

// One tick's block partials: the gather kernel uses at most kMaxStepBlocks
// (4096), the streaming kernel one per 4 waves of its plan.  320 KiB per device.
constexpr uint32_t kScratchParts = 8192;
constexpr int kMaxDevices = 64;

}  // namespace

static __device__ __align__(256) StatPart scratch[kScratchParts];

namespace {

std::mutex g_scratch_mu;
StatPart *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per
// device.  In host code `scratch` names a host-side stand-in, never the
// device memory: only this address goes to kernels.
int scratch_address(StatPart **out)
{
    int dev = -1;
    CAVE_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(CAVE_E_DEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(CAVE_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = static_cast<StatPart *>(p);
    }
    *out = g_scratch_ptr[dev];
    return CAVE_OK;
}

}  // namespace

extern "C" int cave_step_run_static(const CaveSimulationParameters *params, const uint32_t dims[3],
                                    uint32_t kernel, uint32_t *cells_a, uint32_t *cells_b, uint32_t ticks,
                                    CaveTickStats *per_tick, void *stream)
{
    // ---- every argument is checked before the first HIP call ----
    if (!params || !dims || !cells_a || (ticks > 0 && !cells_b))
        return fail(CAVE_E_INVALID, "null argument");
    for (int a = 0; a < 3; a++)
        if (dims[a] == 0 || dims[a] > 65536) return fail(CAVE_E_INVALID, "grid dimensions must be in 1..65536");
    if ((uint64_t)dims[0] * dims[1] * dims[2] > 0xFFFFFFFFull)
        return fail(CAVE_E_INVALID, "grid larger than 2^32 cells: per-tick counters are u32");
    if ((params->shapeAndGridConfiguration & 0x1Fu) != 0u)
        return fail(CAVE_E_UNSUPPORTED, "cave_step_run_static runs the cube shape (id 0) only");
    if (kernel != CAVE_KERNEL_AUTO && kernel != CAVE_KERNEL_GATHER && kernel != CAVE_KERNEL_STREAM_NIBBLE)
        return fail(CAVE_E_INVALID, "kernel must be CAVE_KERNEL_AUTO, _GATHER or _STREAM_NIBBLE");
    const uint32_t flags = uint32_t(params->survivalAndNeighborhoodRules >> 60);
    const bool wrap = (flags & CAVE_WRAP) != 0;
    const uint32_t M = uint32_t(params->birthAndMaxCellStateRules >> 60) + 1u;
    const DevRule rule = make_rule(M, flags, params->birthAndMaxCellStateRules & kPermMask,
                                   params->survivalAndNeighborhoodRules & kPermMask);

    // The reference layout (ceil(X/8) words per row) is the streaming kernel's
    // nibble layout when that pitch is a whole number of 32-cell groups.
    GridGeom g{};
    g.X = dims[0];
    g.Y = dims[1];
    g.Z = dims[2];
    g.pitch = ref_words_per_row(g.X);
    g.planar = 0;
    const bool stream_fits = !wrap && stream_supported(g);
    StreamPlan plan{};
    uint32_t nparts = 0;
    bool use_stream = false;
    if (stream_fits && kernel != CAVE_KERNEL_GATHER) {
        plan = stream_plan(g, 0, 0, 0);
        use_stream = plan.blocks <= kScratchParts;
    }
    if (kernel == CAVE_KERNEL_STREAM_NIBBLE && !use_stream) {
        if (!stream_fits)
            return fail(CAVE_E_UNSUPPORTED, "CAVE_KERNEL_STREAM_NIBBLE needs ceil(X/8) to be a multiple of 4, "
                                            "X <= 2048 and no W flag");
        return fail(CAVE_E_INVALID, "the streaming plan of this grid needs " + std::to_string(plan.blocks) +
                                        " block partials, above the static scratch of " +
                                        std::to_string(kScratchParts) + "; use CAVE_KERNEL_GATHER");
    }
    nparts = use_stream ? plan.blocks : gather_blocks(g);
    if (nparts > kScratchParts)  // gather_blocks is capped at kMaxStepBlocks, below kScratchParts
        return fail(CAVE_E_INVALID, "grid needs more block partials than the static scratch holds");
    if (ticks == 0) return CAVE_OK;

    // ---- set-up: the scratch address on this device (looked up once) ----
    StatPart *parts = nullptr;
    int rc = scratch_address(&parts);
    if (rc) return rc;

    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t *cur = cells_a, *nxt = cells_b;
    for (uint32_t t = 0; t < ticks; t++) {
        rc = use_stream ? launch_stream(cur, nxt, g, rule, plan, parts, st)
                        : launch_gather(cur, nxt, g, rule, parts, st, nullptr, wrap);
        if (rc) return rc;
        if (per_tick && (rc = launch_reduce_ticks(parts, nparts, 1, per_tick + t, st))) return rc;
        std::swap(cur, nxt);
    }
    return CAVE_OK;
}
