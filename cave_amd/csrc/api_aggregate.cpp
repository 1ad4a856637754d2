// api_aggregate.cpp -- cave_aggregate_*: the rule aggregator of a sweep.
//
// Restates RuleAggregator (src/common/ruleAggregator.h:24-135): results are
// grouped by rule -- birth and survival bitmasks, max cell state M,
// neighbourhood flags, shape id -- into run counts per category and the
// minimum / maximum / sum of ticks survived, and written as one JSONL line
// per rule with the reference's key order and its default-ostream floats
// (viability rate and average ticks as float quotients).  Used by the `cave`
// sweep for <prefix>_aggregate.jsonl and by batch callers (one record per
// rule config of a seeded batch).  The reference's map is unordered; lines
// here come in key order, so files compare equal as sets of lines.
#include <cstdint>
#include <fstream>
#include <map>
#include <new>
#include <tuple>

#include "common.h"

using namespace cave;

namespace {

using RuleKey = std::tuple<uint64_t, uint64_t, uint32_t, uint32_t, uint32_t>;  // birth, survival, M, flags, shape

struct RuleStats {
    uint32_t total = 0, viable = 0, possible = 0, unviable = 0;
    uint32_t max_ticks = 0, min_ticks = UINT32_MAX;
    uint64_t sum_ticks = 0;
};

}  // namespace

struct CaveAggregate {
    std::map<RuleKey, RuleStats> rules;
};

namespace {

// ruleAggregator.h:69-90
void record(CaveAggregate *a, const RuleKey &k, uint32_t ticks, uint32_t category)
{
    RuleStats &s = a->rules[k];
    s.total++;
    if (category == 2u) s.viable++;
    else if (category == 1u) s.possible++;
    else s.unviable++;
    if (ticks > s.max_ticks) s.max_ticks = ticks;
    if (ticks < s.min_ticks) s.min_ticks = ticks;
    s.sum_ticks += ticks;
}

}  // namespace

extern "C" {

int cave_aggregate_create(CaveAggregate **out)
{
    if (!out) return fail(CAVE_E_INVALID, "null argument");
    *out = new (std::nothrow) CaveAggregate();
    return *out ? CAVE_OK : fail(CAVE_E_NOMEM, "host allocation failed");
}

int cave_aggregate_record(CaveAggregate *a, const CaveSimulationParameters *params, uint32_t ticks_survived,
                          uint32_t category, uint32_t shape_id)
{
    if (!a || !params) return fail(CAVE_E_INVALID, "null argument");
    if (category > 2u) return fail(CAVE_E_INVALID, "category must be 0 (unviable), 1 (possible) or 2 (viable)");
    // Simulation::DecodeMaxCellState: bits 60-63 of word 0 hold M - 1
    const RuleKey k{params->birthAndMaxCellStateRules & kPermMask, params->survivalAndNeighborhoodRules & kPermMask,
                    (uint32_t)((params->birthAndMaxCellStateRules >> 60) & 0xFu) + 1u,
                    (uint32_t)(params->survivalAndNeighborhoodRules >> 60), shape_id};
    record(a, k, ticks_survived, category);
    return CAVE_OK;
}

int cave_aggregate_record_batch(CaveAggregate *a, const CaveSimSpec *specs, const CaveSimResult *results, size_t n)
{
    if (!a || (n && (!specs || !results))) return fail(CAVE_E_INVALID, "null argument");
    for (size_t i = 0; i < n; i++) {
        const CaveSimSpec &s = specs[i];
        if (results[i].classification > 2u) return fail(CAVE_E_INVALID, "result " + std::to_string(i) + ": bad class");
        const RuleKey k{s.birth & kPermMask, s.survival & kPermMask, s.max_cell_state, s.neighborhood_flags,
                        (s.flags & CAVE_SIM_SHAPE_ERD) ? 1u : 0u};
        record(a, k, results[i].info.numberOfSimulationTicksSurvived, results[i].classification);
    }
    return CAVE_OK;
}

size_t cave_aggregate_rule_count(const CaveAggregate *a) { return a ? a->rules.size() : 0; }

// ruleAggregator.h:105-135
int cave_aggregate_write_jsonl(const CaveAggregate *a, const char *path)
{
    if (!a || !path) return fail(CAVE_E_INVALID, "null argument");
    std::ofstream o(path, std::ios::out | std::ios::trunc);
    if (!o.is_open()) return fail(CAVE_E_INVALID, std::string("cannot open ") + path);
    for (const auto &[k, s] : a->rules) {
        const float avg = s.total ? (float)s.sum_ticks / (float)s.total : 0.0f;
        const float rate = s.total ? (float)s.viable / (float)s.total : 0.0f;
        o << "{\"birth\":" << std::get<0>(k) << ",\"survival\":" << std::get<1>(k)
          << ",\"maxCellState\":" << std::get<2>(k) << ",\"neighborhoodFlags\":" << std::get<3>(k)
          << ",\"shapeId\":" << std::get<4>(k) << ",\"totalRuns\":" << s.total << ",\"viableRuns\":" << s.viable
          << ",\"possibleRuns\":" << s.possible << ",\"unviableRuns\":" << s.unviable
          << ",\"viabilityRate\":" << rate << ",\"minTicksSurvived\":" << (s.min_ticks == UINT32_MAX ? 0 : s.min_ticks)
          << ",\"maxTicksSurvived\":" << s.max_ticks << ",\"avgTicksSurvived\":" << avg << "}\n";
    }
    o.flush();
    if (!o) return fail(CAVE_E_INVALID, std::string("write failed: ") + path);
    return CAVE_OK;
}

int cave_aggregate_destroy(CaveAggregate *a)
{
    delete a;
    return CAVE_OK;
}

}  // extern "C"
