// step_kernels.h -- launch interface of the single-grid step kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "common.h"
#include "device_util.h"

namespace cave {

// Row-pitched nibble layout: cell (x, y, z) lives in word x/8 + pitch*(y + Y*z),
// nibble x%8.  pitch = ceil(X/8) is the reference layout; the streaming kernel
// pads rows to whole 32-cell groups (pitch = 4*ceil(X/32)).
struct GridGeom {
    uint32_t X, Y, Z;
    uint32_t pitch;
    uint32_t planar = 0;  // 1: bit-planar 32-cell groups (pitch = 4 words per group)
};

// Builds the initial cells in the geometry's layout (nibble or planar).
int launch_init(uint32_t *cells, const GridGeom &g, const DevInit &init, uint32_t M, hipStream_t st);
int launch_init_planar(uint32_t *cells, const GridGeom &g, const DevInit &init, uint32_t M, hipStream_t st);
// Layout conversion between the reference nibble layout (ceil(X/8) words per
// row) and the planar layout of g.
int launch_ref_to_planar(const uint32_t *ref, uint32_t *planar, const GridGeom &g, hipStream_t st);
int launch_planar_to_ref(const uint32_t *planar, uint32_t *ref, const GridGeom &g, hipStream_t st);

int launch_sanitize_nibbles(uint32_t *cells, const GridGeom &g, hipStream_t st);
// The gather and ERD kernels are grid-stride loops over the packed words with
// at most this many blocks (16 waves per CU on a 256-CU part), so one tick
// leaves at most this many block partials whatever the grid size.
constexpr uint32_t kMaxStepBlocks = 4096;
uint32_t gather_blocks(const GridGeom &g);
// done (nullable): device flag; when set the launch does nothing (the plain
// per-config driver ends a simulation on the device without a host round trip)
// ERD shape, linear layout (step_erd.hip)
uint32_t erd_blocks(uint32_t X, uint32_t Y, uint32_t Z);
int launch_erd(const uint32_t *in, uint32_t *out, uint32_t X, uint32_t Y, uint32_t Z, const DevRule &rule, bool wrap,
               StatPart *partials, hipStream_t st, const uint32_t *done = nullptr);
int launch_init_erd(uint32_t *cells, uint32_t X, uint32_t Y, uint32_t Z, const DevInit &init, uint32_t M,
                    hipStream_t st);

// wrap: every axis wraps (the render path's W flag)
int launch_gather(const uint32_t *in, uint32_t *out, const GridGeom &g, const DevRule &rule, StatPart *partials,
                  hipStream_t st, const uint32_t *done = nullptr, bool wrap = false);

// One block per tick reduces that tick's block partials into out[tick].  With
// count (device memory, 0) and flag (host-mapped) given, the last block to
// finish sets *flag to 1 once every row is visible to the host.
int launch_reduce_ticks(const StatPart *partials, uint32_t nparts, uint32_t ticks, CaveTickStats *out,
                        hipStream_t st, uint32_t *count = nullptr, uint32_t *flag = nullptr);

// ---- bit-sliced streaming kernel ----
struct StreamPlan {
    uint32_t lpr;      // lanes per x-row (power of two, 2..16)
    uint32_t rpl;      // rows per lane (y register blocking)
    uint32_t tz;       // z planes per wave
    uint32_t waves;    // waves per block
    uint32_t blocks;   // grid size
    uint32_t ytiles;   // y tiles of (64/lpr)*rpl rows
    uint32_t ztiles;
};

bool stream_supported(const GridGeom &g);
// wrap (the W flag): the planar layout, grids of under 1 GiB per buffer
bool stream_wrap_supported(const GridGeom &g);
StreamPlan stream_plan(const GridGeom &g, uint32_t rpl, uint32_t tz, uint32_t waves);
int launch_stream(const uint32_t *in, uint32_t *out, const GridGeom &g, const DevRule &rule, const StreamPlan &p,
                  StatPart *partials, hipStream_t st, const uint32_t *done = nullptr, bool wrap = false);

// ---- plain per-config manager: reduces one tick's partials and applies the
// search manager on the device (searchManager.slang:98-149) ----
struct PlainState {
    CaveGridInfo info;
    uint32_t done;
    uint32_t pad;
    CaveTickStats last;
};
int launch_plain_manager(const StatPart *partials, uint32_t nparts, PlainState *state, uint32_t ticks,
                         uint32_t sim_flags, CaveTickStats *per_tick, uint32_t total_cells, hipStream_t st);

}  // namespace cave
