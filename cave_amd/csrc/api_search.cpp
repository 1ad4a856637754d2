// api_search.cpp -- cave_search_*: one --mode search chunk (SearchSystem).
//
// The reference's SearchSystem (src/vulkan/systems/searchSystem.{h,cpp})
// keeps one GPU slot per (maxCellState, neighbourhood subset) and walks each
// slot's (birth, survival) index space one permutation at a time, one tick
// per host round trip, restarting the cells from the chunk's initial state
// whenever a permutation completes (searchSystem.cpp:496-526).  Every
// evaluation is therefore independent of the others, so this driver
// enumerates the same evaluations (slots of searchMode.cpp:49-102, the
// partitioned survival range of searchSystem.cpp:86-96 / 164-200, the K-capped
// index space of permutationMath.h:101-128) and runs them as batches on the
// persistent engine.  Results are classified (searchSystem.cpp:695-711) and
// streamed to the same three JSONL files with the same keys
// (searchSystem.cpp:959-978), one line per evaluation, flushed per batch.
//
// Differences from the reference, by design: evaluations are emitted in
// enumeration order (the reference emits in completion order), and the
// partitioned ranges are covered exactly once (the reference's work stealing
// can evaluate a permutation twice, benchmarks/reference/debug/
// bench_grand_results_DEBUG.csv:7-8).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "common.h"
#include "host_util.h"

extern "C" int cave_batch_run(const CaveSimSpec *, size_t, CaveSimResult *, CaveBatchOptions *);

using namespace cave;

namespace {

struct Slot {
    uint32_t M, flags;
    std::vector<uint64_t> masks;  // K > 0: valid bitmasks (1-based index space)
    uint64_t max_perm;            // birth index space size
    uint64_t surv_start, surv_end;
};

struct PendingEval {
    uint32_t slot;
    uint64_t birth_idx, surv_idx;
};

}  // namespace

struct CaveSearch {
    CaveSearchSpec spec{};
    std::string folder, prefix;
    std::vector<Slot> slots;
    uint64_t total = 0, completed = 0;
    uint64_t counts[3] = {0, 0, 0};  // unviable, possible, viable
    // enumeration cursor
    uint32_t cur_slot = 0;
    uint64_t cur_surv = 0, cur_birth = 0;
    std::vector<CaveSearchResult> ready;  // not yet drained
    std::ofstream files[3];               // unviable, possible, viable
    std::string paths[3];
    uint32_t spawn[3] = {0, 0, 0};
    double density = 0.0;
};

namespace {

// searchSystem.cpp:226-236 (spawnSizeIndex 0 = the full spawn area)
uint32_t resolve_spawn(uint32_t axis, uint32_t size_index)
{
    if (size_index == 0) return axis;
    const uint32_t scaled = (uint32_t)(((uint64_t)axis * (size_index + 1) + 15) / 16);
    return std::max(1u, std::min(axis, scaled));
}

bool advance(CaveSearch *h, PendingEval &e)
{
    while (h->cur_slot < h->slots.size()) {
        const Slot &s = h->slots[h->cur_slot];
        if (h->cur_surv == 0) {
            h->cur_surv = s.surv_start;
            h->cur_birth = 1;
        }
        if (h->cur_surv > s.surv_end || s.surv_start > s.surv_end) {
            h->cur_slot++;
            h->cur_surv = 0;
            continue;
        }
        e.slot = h->cur_slot;
        e.birth_idx = h->cur_birth;
        e.surv_idx = h->cur_surv;
        // searchManager.slang:56-96: birth index fastest, then survival
        if (++h->cur_birth > s.max_perm) {
            h->cur_birth = 1;
            h->cur_surv++;
        }
        return true;
    }
    return false;
}

uint64_t mask_of(const Slot &s, uint64_t idx) { return s.masks.empty() ? idx : s.masks[idx - 1]; }

// nlohmann::json's default object is a std::map: keys come out sorted, and
// dump() writes unsigned integers in decimal with no spaces.
std::string jsonl_line(const CaveSearchResult &r)
{
    char buf[512];
    std::snprintf(buf, sizeof buf,
                  "{\"birthAndMaxCellStateRules\":%llu,\"finalAliveCount\":%u,\"maxAliveCount\":%u,"
                  "\"shapeAndGridConfiguration\":%llu,\"survivalAndNeighborhoodRules\":%llu,\"ticksSurvived\":%u,"
                  "\"totalBirths\":%u,\"totalDeaths\":%u}",
                  (unsigned long long)r.permutation.birthAndMaxCellStateRules, r.finalAliveCount, r.maxAliveCount,
                  (unsigned long long)r.permutation.shapeAndGridConfiguration,
                  (unsigned long long)r.permutation.survivalAndNeighborhoodRules, r.ticksSurvived, r.totalBirths,
                  r.totalDeaths);
    return buf;
}

void release(CaveSearch *h)
{
    for (auto &f : h->files)
        if (f.is_open()) f.close();
    delete h;
}

}  // namespace

extern "C" {

int cave_search_create(const CaveSearchSpec *spec, CaveSearch **out)
{
    if (!spec || !out) return fail(CAVE_E_INVALID, "null argument");
    *out = nullptr;
    const CaveChunkConfiguration &c = spec->chunk;
    if (c.shapeId > 1) return fail(CAVE_E_UNSUPPORTED, "shape ids are 0 (cube) and 1 (ERD)");
    if (spec->neighborhood_flags & ~15u) return fail(CAVE_E_INVALID, "bad neighbourhood flags");
    if ((spec->neighborhood_flags & 7u) == 0) return fail(CAVE_E_INVALID, "neighbourhood must include F, E or C");
    if (c.gridDimensionX == 0 || c.gridDimensionY == 0 || c.gridDimensionZ == 0)
        return fail(CAVE_E_INVALID, "grid dimensions must be >= 1");
    if (spec->min_max_cell_state < 1 || spec->max_max_cell_state > 16)
        return fail(CAVE_E_INVALID, "max cell state range must lie in 1..16");
    if (spec->max_ticks == 0) return fail(CAVE_E_INVALID, "ticks must be >= 1");
    if (c.spawnRegionShapeId > 1) return fail(CAVE_E_INVALID, "spawn region must be random (0) or filled (1)");
    if (c.spawnDensityIndex > 15) return fail(CAVE_E_INVALID, "density index must be 0..15");
    const uint32_t P = std::max(1u, c.partitionCount);
    if (c.partitionIndex >= P) return fail(CAVE_E_INVALID, "partition index out of range");

    CaveSearch *h = new (std::nothrow) CaveSearch();
    if (!h) return fail(CAVE_E_NOMEM, "host allocation failed");
    h->spec = *spec;
    h->folder = spec->output_folder ? spec->output_folder : "";
    h->prefix = spec->file_prefix ? spec->file_prefix : "";
    h->spec.output_folder = nullptr;
    h->spec.file_prefix = nullptr;
    const uint32_t dims[3] = {c.gridDimensionX, c.gridDimensionY, c.gridDimensionZ};
    const uint32_t area[3] = {c.spawnAreaDimensionX, c.spawnAreaDimensionY, c.spawnAreaDimensionZ};
    for (int a = 0; a < 3; a++) h->spawn[a] = resolve_spawn(std::min(area[a], dims[a]), c.spawnSizeIndex);
    h->density = (double)((float)c.spawnDensityIndex + 1.0f) / 16.0;  // searchSystem.cpp:240

    // slots: searchMode.cpp:49-102 -- every non-empty subset of the selected
    // F / E / C, each also with W when W is selected (W wraps on ERD; the cube
    // search kernel ignores it, searchSimulationCube.slang:26-29)
    const uint32_t sel = spec->neighborhood_flags & 7u;
    const bool erd = c.shapeId == 1u;
    std::vector<uint32_t> configs;
    for (uint32_t subset = 1; subset <= sel; subset++) {
        if ((subset & sel) != subset) continue;
        configs.push_back(subset);
        if (spec->neighborhood_flags & CAVE_WRAP) configs.push_back(subset | CAVE_WRAP);
    }
    for (int M = spec->min_max_cell_state; M <= spec->max_max_cell_state; M++)
        for (uint32_t subset : configs) {
            Slot s;
            s.M = (uint32_t)M;
            s.flags = subset;
            // neighbour counts per shape (src/shapes/*.h): ERD has 12 faces, 4 edges, no corners
            const uint32_t n = erd ? ((subset & 1u) ? 12u : 0u) + ((subset & 2u) ? 4u : 0u) : neighbor_count(subset);
            if (n == 0) continue;  // no neighbours: no rule space
            if (spec->max_rule_bits > 0) {
                s.masks = enumerate_valid_bitmasks(n, spec->max_rule_bits);
                s.max_perm = s.masks.size();
            } else {
                s.max_perm = n < 60 ? ((uint64_t(1) << n) - 1) : kPermMask;
            }
            // searchSystem.cpp:86-96, 164-180: partition i of P
            s.surv_start = 1;
            s.surv_end = s.max_perm;
            if (P > 1) {
                s.surv_start = 1 + (uint64_t)c.partitionIndex * s.max_perm / P;
                s.surv_end = (uint64_t)(c.partitionIndex + 1) * s.max_perm / P;
            }
            if (s.surv_start <= s.surv_end) h->total += (s.surv_end - s.surv_start + 1) * s.max_perm;
            h->slots.push_back(std::move(s));
        }

    if (!h->folder.empty()) {
        static const char *names[3] = {"_unviable.jsonl", "_possible.jsonl", "_viable.jsonl"};
        for (int k = 0; k < 3; k++) {
            h->paths[k] = h->folder + h->prefix + names[k];
            h->files[k].open(h->paths[k], std::ios::out | std::ios::trunc);
            if (!h->files[k].is_open()) {
                const std::string p = h->paths[k];
                release(h);
                return fail(CAVE_E_INVALID, "cannot open " + p + " (does the folder exist?)");
            }
        }
    }
    *out = h;
    return CAVE_OK;
}

// PermutationMath::EstimateChunkVramBytes (permutationMath.h:144-153): three
// packed cell buffers of ceil(V/8) words and ~128 B of small buffers per slot,
// plus ~16 MiB of pipeline and descriptor objects per chunk.
uint64_t cave_estimate_chunk_vram_bytes(uint32_t X, uint32_t Y, uint32_t Z, uint32_t slot_count)
{
    const uint64_t packed = ((uint64_t)X * Y * Z + 7u) / 8u;
    return (3u * 4u * packed + 128u) * slot_count + (16ull << 20);
}

int cave_search_step(CaveSearch *h, uint64_t max_evals) { return cave_search_step_group(&h, 1, max_evals); }

int cave_search_step_group(CaveSearch *const *chunks, size_t n_chunks, uint64_t max_evals)
{
    if (!chunks && n_chunks) return fail(CAVE_E_INVALID, "null chunk list");
    for (size_t c = 0; c < n_chunks; c++)
        if (!chunks[c]) return fail(CAVE_E_INVALID, "null handle");
    const uint64_t kBatch = 1u << 20;
    uint64_t budget = max_evals ? max_evals : ~0ull;
    std::vector<CaveSimSpec> specs;
    std::vector<PendingEval> evals;
    std::vector<uint32_t> owner;  // chunk of each evaluation
    size_t cur = 0;               // chunks before `cur` have nothing left to enumerate
    while (budget > 0) {
        specs.clear();
        evals.clear();
        owner.clear();
        PendingEval e{};
        while (specs.size() < std::min<uint64_t>(kBatch, budget) && cur < n_chunks) {
            CaveSearch *h = chunks[cur];
            if (!advance(h, e)) {
                cur++;
                continue;
            }
            owner.push_back((uint32_t)cur);
            const Slot &s = h->slots[e.slot];
            CaveSimSpec sp{};
            sp.dims[0] = h->spec.chunk.gridDimensionX;
            sp.dims[1] = h->spec.chunk.gridDimensionY;
            sp.dims[2] = h->spec.chunk.gridDimensionZ;
            sp.max_cell_state = s.M;
            sp.neighborhood_flags = s.flags;
            sp.ticks = h->spec.max_ticks;
            sp.birth = mask_of(s, e.birth_idx);
            sp.survival = mask_of(s, e.surv_idx);
            for (int a = 0; a < 3; a++) sp.spawn[a] = h->spawn[a];
            sp.spawn_mode = h->spec.chunk.spawnRegionShapeId == 0 ? CAVE_SPAWN_RANDOM : CAVE_SPAWN_FILLED;
            sp.density = h->density;
            sp.seed = h->spec.chunk.runSeedIndex;
            sp.flags = h->spec.sim_flags & ~(uint32_t)(CAVE_SIM_FINAL_HASH | CAVE_SIM_SHAPE_ERD);
            if (h->spec.chunk.shapeId == 1u) sp.flags |= CAVE_SIM_SHAPE_ERD;
            specs.push_back(sp);
            evals.push_back(e);
        }
        if (specs.empty()) break;
        std::vector<CaveSimResult> res(specs.size());
        CaveBatchOptions opt{};
        opt.engine = CAVE_ENGINE_PERSISTENT;
        int rc = cave_batch_run(specs.data(), specs.size(), res.data(), &opt);
        if (rc) return rc;
        for (size_t i = 0; i < specs.size(); i++) {
            CaveSearch *h = chunks[owner[i]];
            const CaveSimSpec &sp = specs[i];
            CaveSearchResult r{};
            // bitmask form (searchSystem.cpp:668-685)
            r.permutation.birthAndMaxCellStateRules = ((uint64_t)((sp.max_cell_state - 1u) & 0xFu) << 60) | sp.birth;
            r.permutation.survivalAndNeighborhoodRules = ((uint64_t)sp.neighborhood_flags << 60) | sp.survival;
            // the reference's chunks carry shape id 0 whatever the shader's shape
            // (searchMode.cpp:197, searchSystem.cpp:198-199), so its records hold 0
            r.permutation.shapeAndGridConfiguration = 0;
            r.ticksSurvived = res[i].info.numberOfSimulationTicksSurvived;
            r.totalBirths = res[i].info.totalNumberOfCellsThatWereBorn;
            r.totalDeaths = res[i].info.totalNumberOfCellsThatDied;
            r.maxAliveCount = res[i].info.maxAliveCount;
            r.finalAliveCount = res[i].info.finalAliveCount;
            r.category = res[i].classification;
            h->counts[r.category]++;
            h->completed++;
            if (h->files[r.category].is_open()) h->files[r.category] << jsonl_line(r) << '\n';
            h->ready.push_back(r);
        }
        for (size_t c = 0; c < n_chunks; c++)
            for (auto &f : chunks[c]->files)
                if (f.is_open()) f.flush();
        budget -= specs.size();
    }
    return CAVE_OK;
}

int cave_search_is_running(const CaveSearch *h) { return h && h->completed < h->total ? 1 : 0; }

int cave_search_progress(const CaveSearch *h, CaveSearchProgress *out)
{
    if (!h || !out) return fail(CAVE_E_INVALID, "null argument");
    out->total = h->total;
    out->completed = h->completed;
    out->unviable = h->counts[0];
    out->possible = h->counts[1];
    out->viable = h->counts[2];
    return CAVE_OK;
}

int cave_search_drain(CaveSearch *h, CaveSearchResult *out, size_t cap, size_t *n)
{
    if (!h || !n) return fail(CAVE_E_INVALID, "null argument");
    const size_t k = out ? std::min(cap, h->ready.size()) : 0;
    if (k) std::copy_n(h->ready.begin(), k, out);
    h->ready.erase(h->ready.begin(), h->ready.begin() + (ptrdiff_t)k);
    *n = k;
    return CAVE_OK;
}

int cave_search_write_summary(CaveSearch *h, char *msg, size_t cap)
{
    if (!h) return fail(CAVE_E_INVALID, "null handle");
    for (auto &f : h->files)
        if (f.is_open()) f.flush();
    // searchSystem.cpp:989-991 (the benchmark scripts parse this line)
    char buf[4096];
    std::snprintf(buf, sizeof buf,
                  "Search results flushed \xe2\x80\x94 viable: %llu, possible: %llu, unviable: %llu (streamed to %s, %s, %s)",
                  (unsigned long long)h->counts[2], (unsigned long long)h->counts[1],
                  (unsigned long long)h->counts[0], h->paths[2].c_str(), h->paths[1].c_str(), h->paths[0].c_str());
    if (msg && cap) {
        std::strncpy(msg, buf, cap - 1);
        msg[cap - 1] = 0;
    }
    return CAVE_OK;
}

int cave_search_destroy(CaveSearch *h)
{
    if (h) release(h);
    return CAVE_OK;
}

}  // extern "C"
