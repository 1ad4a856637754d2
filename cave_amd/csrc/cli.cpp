// cli.cpp -- the `cave` command line: --mode search (the reference's search
// sweep, src/modes/searchMode.cpp) and --mode step (the headless step loop),
// on top of the C-ABI.  Flag names, defaults, validation messages and exit
// codes follow src/main.cpp:21-221, 343-377: 0 ok, 1 validation error,
// 2 argument parse error, 4 exception.
#include <algorithm>
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <filesystem>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/cave.h"

namespace {

struct ParseError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// --name value / --name=value / boolean --name
class Args {
public:
    Args(int argc, char **argv)
    {
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a == "-h") a = "--help";
            if (a.rfind("--", 0) != 0) throw ParseError("unexpected argument '" + a + "'");
            a = a.substr(2);
            const size_t eq = a.find('=');
            if (eq != std::string::npos) {
                kv_[a.substr(0, eq)] = a.substr(eq + 1);
            } else if (is_bool(a)) {
                kv_[a] = "true";
            } else {
                if (i + 1 >= argc) throw ParseError("option '--" + a + "' needs a value");
                kv_[a] = argv[++i];
            }
        }
    }
    bool has(const std::string &k) const { return kv_.count(k) > 0; }
    std::string str(const std::string &k, const std::string &def) const
    {
        auto it = kv_.find(k);
        return it == kv_.end() ? def : it->second;
    }
    int integer(const std::string &k, int def) const
    {
        auto it = kv_.find(k);
        if (it == kv_.end()) return def;
        char *end = nullptr;
        const long v = std::strtol(it->second.c_str(), &end, 10);
        if (!end || *end) throw ParseError("option '--" + k + "' expects an integer");
        return (int)v;
    }
    double real(const std::string &k, double def) const
    {
        auto it = kv_.find(k);
        if (it == kv_.end()) return def;
        char *end = nullptr;
        const double v = std::strtod(it->second.c_str(), &end);
        if (!end || *end) throw ParseError("option '--" + k + "' expects a number");
        return v;
    }
    bool flag(const std::string &k) const
    {
        auto it = kv_.find(k);
        return it != kv_.end() && it->second != "false" && it->second != "0";
    }

private:
    static bool is_bool(const std::string &k)
    {
        static const char *b[] = {"headless", "dual-gpu", "help", "log-file", "benchmark", "compute-skip",
                                  "legacy-death", "list-gpus"};
        for (const char *x : b)
            if (k == x) return true;
        return false;
    }
    std::map<std::string, std::string> kv_;
};

void log_info(const std::string &m) { std::printf("[info] %s\n", m.c_str()), std::fflush(stdout); }
void log_error(const std::string &m) { std::fprintf(stderr, "[error] %s\n", m.c_str()); }

const char *kHelp =
    "cave -- MI355X-native CAVE step and rule search\n"
    "  --mode search|step       application mode\n"
    "  search: --grid N | --grid-sweep min:max:step, --neighborhood FEC, --max-rule-bits K,\n"
    "          --min-cs M0 --max-cs-range M1, --ticks T, --chunks-per-config P, --spawn S,\n"
    "          --spawn-mode random|filled, --spawn-density d | --density-sweep a:b:s,\n"
    "          --seed-runs R, --output DIR, --legacy-death\n"
    "          --shape cube|ed (W in --neighborhood adds wrapping slots)\n"
    "          --gpu-vram-budget-gb G (default 40): chunks admitted together under the reference's\n"
    "          VRAM estimate share launches; a chunk above it alone aborts the sweep\n"
    "          (accepted and ignored: --headless --dual-gpu --search-workgroup-size)\n"
    "  step:   --grid N, --birth \"5,6\", --survival \"4-6\", --max-cs M, --neighborhood FE[W], --frames T,\n"
    "          --spawn S, --spawn-mode random|filled, --spawn-density d, --seed s, --dump FILE\n"
    "          (--mode render --headless: the same loop with the render path's unsalted random spawn;\n"
    "           W in --neighborhood wraps every axis)\n";

uint32_t parse_flags(const std::string &nb, bool allow_wrap)
{
    uint32_t f = 0;
    if (nb.find('F') != std::string::npos) f |= CAVE_FACES;
    if (nb.find('E') != std::string::npos) f |= CAVE_EDGES;
    if (nb.find('C') != std::string::npos) f |= CAVE_CORNERS;
    if (allow_wrap && nb.find('W') != std::string::npos) f |= CAVE_WRAP;
    return f;
}

int check(int rc)
{
    if (rc != CAVE_OK) throw std::runtime_error(std::string("cave error ") + std::to_string(rc) + ": " + cave_last_error());
    return rc;
}

int run_search(const Args &a)
{
    // src/main.cpp:248-258: cube, or ed / elongateddodecahedron (case-insensitive)
    std::string shape = a.str("shape", "cube");
    std::transform(shape.begin(), shape.end(), shape.begin(), [](unsigned char ch) { return (char)std::tolower(ch); });
    uint32_t shape_id = 0;
    if (shape == "cube") shape_id = 0;
    else if (shape == "ed" || shape == "elongateddodecahedron") shape_id = 1;
    else {
        log_error("Unknown --shape: " + a.str("shape", "") + ". Use 'cube' or 'ed'.");
        return 1;
    }
    const std::string nb = a.str("neighborhood", "F");
    const uint32_t sel = parse_flags(nb, true);
    if (!(sel & 7u)) {
        log_error("--neighborhood must include at least one of F / E / C (got \"" + nb + "\")");
        return 1;
    }
    const int min_cs = a.integer("min-cs", 3), max_cs = a.integer("max-cs-range", 10);
    const int ticks = a.integer("ticks", 50);
    const int grid = a.integer("grid", 49), spawn = a.integer("spawn", 11);
    const std::string spawn_mode = a.str("spawn-mode", "filled");
    if (spawn_mode != "random" && spawn_mode != "filled") {
        log_error("Unknown spawn-mode: " + spawn_mode + ". Use random or filled.");
        return 1;
    }
    const double density = std::clamp(a.real("spawn-density", 0.5), 0.0, 1.0);
    const int seed_runs = std::max(1, a.integer("seed-runs", 1));
    const int K = std::max(0, a.integer("max-rule-bits", 0));
    const int P = std::max(1, a.integer("chunks-per-config", 64));
    const int wgs = a.integer("search-workgroup-size", 256);
    if (wgs != 32 && wgs != 64 && wgs != 128 && wgs != 256 && wgs != 512) {
        log_error("--search-workgroup-size must be one of {32, 64, 128, 256, 512} (got " + std::to_string(wgs) + ")");
        return 1;
    }
    if (min_cs < 1 || max_cs > 16 || ticks < 1 || grid < 1) {
        log_error("invalid --min-cs / --max-cs-range / --ticks / --grid");
        return 1;
    }
    std::vector<uint32_t> grids;
    if (a.has("grid-sweep")) {
        int g0 = 0, g1 = 0, gs = 0;
        if (std::sscanf(a.str("grid-sweep", "").c_str(), "%d:%d:%d", &g0, &g1, &gs) != 3 || g0 <= 0 || g1 < g0 ||
            gs <= 0) {
            log_error("--grid-sweep must be min:max:step with positive ints and max >= min");
            return 1;
        }
        for (int g = g0; g <= g1; g += gs) grids.push_back((uint32_t)g);
    } else {
        grids.push_back((uint32_t)grid);
    }
    const bool random = spawn_mode == "random";
    const uint32_t base_density_index = (uint32_t)std::clamp((int)(density * 16.0f) - 1, 0, 15);  // searchMode.cpp:135-136
    std::vector<uint32_t> densities;
    if (a.has("density-sweep") && random) {
        float d0 = 0, d1 = 0, ds = 0;
        if (std::sscanf(a.str("density-sweep", "").c_str(), "%f:%f:%f", &d0, &d1, &ds) != 3 || d0 <= 0 || d1 < d0 ||
            ds <= 0 || d0 > 1 || d1 > 1) {
            log_error("--density-sweep must be min:max:step in (0,1] with max >= min");
            return 1;
        }
        for (float d = d0; d <= d1 + 1e-6f; d += ds) densities.push_back((uint32_t)std::clamp((int)(d * 16.0f) - 1, 0, 15));
    } else {
        densities.push_back(base_density_index);
    }

    // searchMode.cpp:104-123: output folder and file prefix
    std::string tag;
    if (sel & CAVE_FACES) tag += "F";
    if (sel & CAVE_EDGES) tag += "E";
    if (sel & CAVE_CORNERS) tag += "C";
    if (sel & CAVE_WRAP) tag += "W";
    std::string base = "cs" + std::to_string(min_cs) + "-" + std::to_string(max_cs) + "_" + tag + "_t" + std::to_string(ticks);
    if (K > 0) base += "_b" + std::to_string(K);
    std::string out = a.str("output", "data/search/");
    if (!out.empty() && out.back() != '/') out += "/";
    // searchMode.cpp:120: <output><shape name>/
    const std::string folder = out + (shape_id == 1 ? "ElongatedRhombicDodecahedron/" : "cube/");
    std::filesystem::create_directories(folder);
    const uint32_t sim_flags = a.flag("legacy-death") ? CAVE_SIM_LEGACY_DEATH : 0u;

    // the reference's slot list, which its VRAM estimate counts: every M of
    // the range x every non-empty subset of the selected F / E / C, each again
    // with W when W is selected (searchMode.cpp:49-102)
    uint32_t n_subsets = 0;
    for (uint32_t subset = 1; subset <= (sel & 7u); subset++)
        if ((subset & sel) == subset) n_subsets++;
    const uint32_t slot_count = (uint32_t)(max_cs - min_cs + 1) * n_subsets * ((sel & CAVE_WRAP) ? 2u : 1u);
    // searchMode.cpp:227-229 (float budget, at least 0.5 GB)
    const uint64_t budget =
        (uint64_t)(std::max(0.5f, (float)a.real("gpu-vram-budget-gb", 40.0)) * 1024.0 * 1024.0 * 1024.0);

    struct Chunk {
        uint32_t grid, density;
        int seed, part;
    };
    std::vector<Chunk> chunks;
    for (uint32_t g : grids)
        for (uint32_t d : densities)
            for (int s = 0; s < seed_runs; s++)
                for (int p = 0; p < P; p++) chunks.push_back({g, d, s, p});
    char head[512];
    std::snprintf(head, sizeof head, "Starting search sweep: %zu chunk(s), maxCS %d-%d, VRAM budget %.1f GB/GPU, output: %s",
                  chunks.size(), min_cs, max_cs, budget / 1073741824.0, folder.c_str());
    log_info(head);
    // RuleAggregator (src/common/ruleAggregator.h:24-135), in the library
    CaveAggregate *agg = nullptr;
    check(cave_aggregate_create(&agg));
    uint64_t tot[3] = {0, 0, 0};
    const auto t0 = std::chrono::steady_clock::now();
    // The reference keeps as many chunks in flight as its VRAM estimate
    // admits (searchMode.cpp:349-404); here the admitted chunks share batched
    // launches.  A chunk that alone exceeds the budget aborts the rest of the
    // sweep with the reference's error (searchMode.cpp:358-366).
    size_t i = 0;
    while (i < chunks.size()) {
        size_t j = i;
        uint64_t used = 0;
        bool abort_sweep = false;
        while (j < chunks.size()) {
            const uint32_t g = chunks[j].grid;
            const uint64_t est = cave_estimate_chunk_vram_bytes(g, g, g, slot_count);
            if (est > budget) {
                if (j == i) {
                    char m[256];
                    std::snprintf(m, sizeof m,
                                  "Chunk requires %.2f GB but VRAM budget is only %.2f GB per GPU. Aborting sweep.",
                                  est / 1073741824.0, budget / 1073741824.0);
                    log_error(m);
                    abort_sweep = true;
                }
                break;
            }
            if (used + est > budget) break;
            used += est;
            j++;
        }
        if (abort_sweep) break;
        std::vector<CaveSearch *> hs;
        std::vector<std::string> prefixes;
        prefixes.reserve(j - i);
        for (size_t k = i; k < j; k++) {
            const Chunk &ck = chunks[k];
            CaveSearchSpec spec{};
            CaveChunkConfiguration &c = spec.chunk;
            c.shapeId = shape_id;
            c.gridDimensionX = c.gridDimensionY = c.gridDimensionZ = ck.grid;
            c.spawnAreaDimensionX = c.spawnAreaDimensionY = c.spawnAreaDimensionZ = std::min((uint32_t)spawn, ck.grid);
            c.spawnRegionShapeId = random ? 0u : 1u;
            c.spawnDensityIndex = ck.density;
            c.runSeedIndex = (uint32_t)ck.seed;
            c.partitionIndex = (uint32_t)ck.part;
            c.partitionCount = (uint32_t)P;
            spec.neighborhood_flags = sel;
            spec.max_rule_bits = (uint32_t)K;
            spec.min_max_cell_state = min_cs;
            spec.max_max_cell_state = max_cs;
            spec.max_ticks = (uint32_t)ticks;
            spec.sim_flags = sim_flags;
            const std::string gs = std::to_string(ck.grid);
            prefixes.push_back(base + "_c" + std::to_string(k + 1) + "_g" + gs + "x" + gs + "x" + gs + "_d" +
                               std::to_string(ck.density) + "_s" + std::to_string(ck.seed) + "_gpu0");
            spec.output_folder = folder.c_str();
            spec.file_prefix = prefixes.back().c_str();
            CaveSearch *h = nullptr;
            const int rc = cave_search_create(&spec, &h);
            if (rc != CAVE_OK) {
                for (CaveSearch *x : hs) cave_search_destroy(x);
                check(rc);
            }
            hs.push_back(h);
        }
        char m[256];
        std::snprintf(m, sizeof m, "Running chunks %zu-%zu of %zu together (estimated VRAM %.2f/%.2f GB)", i + 1, j,
                      chunks.size(), used / 1073741824.0, budget / 1073741824.0);
        log_info(m);
        const int rc = cave_search_step_group(hs.data(), hs.size(), 0);
        if (rc != CAVE_OK) {
            for (CaveSearch *x : hs) cave_search_destroy(x);
            check(rc);
        }
        for (CaveSearch *h : hs) {
            std::vector<CaveSearchResult> res(4096);
            size_t n = 0;
            do {
                check(cave_search_drain(h, res.data(), res.size(), &n));
                for (size_t r = 0; r < n; r++)
                    check(cave_aggregate_record(agg, &res[r].permutation, res[r].ticksSurvived, res[r].category,
                                                (uint32_t)(res[r].permutation.shapeAndGridConfiguration & 0x1Fu)));
            } while (n == res.size());
            CaveSearchProgress pr{};
            check(cave_search_progress(h, &pr));
            tot[0] += pr.viable;
            tot[1] += pr.possible;
            tot[2] += pr.unviable;
            char msg[4096];
            check(cave_search_write_summary(h, msg, sizeof msg));
            log_info(msg);
            cave_search_destroy(h);
        }
        i = j;
    }
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    log_info("Sweep complete. Viable: " + std::to_string(tot[0]) + ", Possible: " + std::to_string(tot[1]) +
             ", Unviable: " + std::to_string(tot[2]));
    const std::string agg_path = folder + base + "_aggregate.jsonl";
    check(cave_aggregate_write_jsonl(agg, agg_path.c_str()));
    log_info("Aggregated " + std::to_string(cave_aggregate_rule_count(agg)) + " unique rules to " + agg_path);
    cave_aggregate_destroy(agg);
    char line[256];
    std::snprintf(line, sizeof line, "Evaluated %llu rule configs in %.3f s (%.0f configs/s)",
                  (unsigned long long)(tot[0] + tot[1] + tot[2]), secs, (tot[0] + tot[1] + tot[2]) / std::max(secs, 1e-9));
    log_info(line);
    return 0;
}

// --mode step, or --mode render --headless: the headless step loop
// (renderingMode.cpp:166-343).  The render mode's random spawn is
// cellInit.slang's (no seed salt, float threshold); W in --neighborhood wraps.
int run_step(const Args &a, bool render)
{
    const std::string nb = a.str("neighborhood", "F");
    const uint32_t flags = parse_flags(nb, true);
    if (!(flags & 7u)) {
        log_error("--neighborhood must include at least one of F / E / C (got \"" + nb + "\")");
        return 1;
    }
    const int grid = a.integer("grid", 49), frames = a.integer("frames", 20), M = a.integer("max-cs", 5);
    const int spawn = a.integer("spawn", 11);
    const std::string spawn_mode = a.str("spawn-mode", "filled");
    if (spawn_mode != "random" && spawn_mode != "filled") {
        log_error("Unknown spawn-mode: " + spawn_mode + ". Use random or filled.");
        return 1;
    }
    if (M < 1 || M > 16 || grid < 1 || frames < 0) {
        log_error("invalid --max-cs / --grid / --frames");
        return 1;
    }
    CaveStepSpec spec{};
    spec.dims[0] = spec.dims[1] = spec.dims[2] = (uint32_t)grid;
    const uint64_t B = cave_rule_string_to_bitmask(a.str("birth", "").c_str());
    const uint64_t S = cave_rule_string_to_bitmask(a.str("survival", "").c_str());
    spec.params.birthAndMaxCellStateRules = ((uint64_t)((M - 1) & 0xF) << 60) | B;  // guiSystem.cpp:370-388
    spec.params.survivalAndNeighborhoodRules = ((uint64_t)flags << 60) | S;
    spec.spawn[0] = spec.spawn[1] = spec.spawn[2] = (uint32_t)std::min(spawn, grid);
    spec.spawn_mode = spawn_mode == "random" ? (render ? CAVE_SPAWN_RENDER_RANDOM : CAVE_SPAWN_RANDOM)
                                             : CAVE_SPAWN_FILLED;
    spec.density = std::clamp(a.real("spawn-density", 0.5), 0.0, 1.0);
    spec.seed = (uint32_t)a.integer("seed", 0);
    CaveStep *h = nullptr;
    check(cave_step_create(&spec, &h));
    std::vector<CaveTickStats> st((size_t)std::max(frames, 1));
    double ms = 0;
    check(cave_step_run(h, (uint32_t)frames, st.data(), &ms));
    for (int t = 0; t < frames; t++) {
        const CaveTickStats &s = st[t];
        std::printf("tick %d dead %u dying %u born %u surviving %u live %u changed %u bbox [%d,%d,%d]-[%d,%d,%d]\n",
                    t + 1, s.snap.numberOfDeadCells, s.snap.numberOfDyingCells, s.snap.numberOfCellsBorn,
                    s.snap.numberOfSurvivingCells, s.live, s.changed, s.bbox_min[0], s.bbox_min[1], s.bbox_min[2],
                    s.bbox_max[0], s.bbox_max[1], s.bbox_max[2]);
    }
    const double cells = (double)grid * grid * grid * frames;
    char line[256];
    std::snprintf(line, sizeof line, "%d ticks of %d^3 in %.3f ms device time (%.3g cell-updates/s)", frames, grid, ms,
                  ms > 0 ? cells / (ms / 1e3) : 0.0);
    log_info(line);
    if (a.has("dump")) {
        std::vector<uint32_t> cellsv(cave_cell_words((uint32_t)grid, (uint32_t)grid, (uint32_t)grid));
        check(cave_step_download(h, cellsv.data(), cellsv.size()));
        std::ofstream f(a.str("dump", ""), std::ios::binary);
        f.write(reinterpret_cast<const char *>(cellsv.data()), (std::streamsize)(cellsv.size() * 4));
    }
    cave_step_destroy(h);
    return 0;
}

}  // namespace

int main(int argc, char **argv)
{
    try {
        Args a(argc, argv);
        if (a.flag("help")) {
            std::fputs(kHelp, stdout);
            return 0;
        }
        const std::string mode = a.str("mode", "");
        if (mode == "search") return run_search(a);
        if (mode == "step" || (mode == "render" && a.flag("headless"))) return run_step(a, mode == "render");
        log_error("Unknown mode: " + mode + ". Use search, step or render --headless (the interactive render, video and looking-glass modes are not implemented here).");
        return 1;
    } catch (const ParseError &e) {
        std::fprintf(stderr, "Argument parse error: %s\n", e.what());
        return 2;
    } catch (const std::exception &e) {
        log_error(std::string("Unhandled exception: ") + e.what());
        return 4;
    }
}
