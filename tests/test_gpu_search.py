"""--mode search on the GPU (cave_search_* and the `cave` CLI) vs the CPU oracle
and the reference's published sweep counts."""
import json
import os
import re
import subprocess

import numpy as np
import pytest

import cave_amd as cv
import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "cave_amd", "_lib", "cave")


def results_key(results):
    return sorted((r.permutation.birthAndMaxCellStateRules, r.permutation.survivalAndNeighborhoodRules,
                   r.ticksSurvived, r.totalBirths, r.totalDeaths, r.maxAliveCount, r.finalAliveCount, r.category)
                  for r in results)


def test_small_sweep_matches_oracle_per_evaluation(tmp_path):
    grid, ticks = 9, 8
    s = cv.Search(grid, cv.FACES | cv.EDGES, 2, 1, 2, ticks, spawn=3, spawn_mode=cv.SPAWN_FILLED,
                  output_folder=str(tmp_path) + "/", file_prefix="t")
    s.run()
    res = s.drain()
    prog = s.progress()
    assert prog["completed"] == prog["total"] == len(res) == 2 * (21 ** 2 + 78 ** 2 + 171 ** 2)
    inits = {M: O.init_cells(grid, grid, grid, M, (3, 3, 3), 1, 0.5, 0) for M in (1, 2)}
    rng = np.random.default_rng(0)
    for i in rng.choice(len(res), 300, replace=False):
        r = res[i]
        M = ((r.permutation.birthAndMaxCellStateRules >> 60) & 0xF) + 1
        flags = r.permutation.survivalAndNeighborhoodRules >> 60
        b = r.permutation.birthAndMaxCellStateRules & ((1 << 60) - 1)
        sv = r.permutation.survivalAndNeighborhoodRules & ((1 << 60) - 1)
        t, info, _, _ = O.run_rule(grid, grid, grid, M, flags, b, sv, ticks, inits[M])
        assert (r.ticksSurvived, r.totalBirths, r.totalDeaths, r.maxAliveCount, r.finalAliveCount) == (
            info.numberOfSimulationTicksSurvived, info.totalNumberOfCellsThatWereBorn,
            info.totalNumberOfCellsThatDied, info.maxAliveCount, info.finalAliveCount)
        assert r.category == O.lib().orc_classify(info, M)
    # JSONL files: one line per evaluation, in the file of its category, keys sorted
    lines = {}
    for cat, name in ((2, "viable"), (1, "possible"), (0, "unviable")):
        with open(tmp_path / f"t_{name}.jsonl") as f:
            lines[cat] = [json.loads(l) for l in f]
        assert len(lines[cat]) == sum(1 for r in res if r.category == cat)
        for l in lines[cat][:5]:
            assert list(l) == sorted(l)
    first = open(tmp_path / "t_viable.jsonl").readline().strip()
    assert re.fullmatch(r'\{"birthAndMaxCellStateRules":\d+,"finalAliveCount":\d+,"maxAliveCount":\d+,'
                        r'"shapeAndGridConfiguration":0,"survivalAndNeighborhoodRules":\d+,"ticksSurvived":\d+,'
                        r'"totalBirths":\d+,"totalDeaths":\d+\}', first)
    summary = s.summary()
    assert re.search(r"viable: (\d+), possible: (\d+), unviable: (\d+)", summary)


def test_partitions_cover_each_permutation_exactly_once():
    base = None
    for P in (1, 3, 7, 64, 500):
        allr = []
        for i in range(P):
            s = cv.Search(7, cv.FACES | cv.CORNERS, 2, 2, 2, 6, spawn=3, partition_index=i, partition_count=P)
            s.run()
            allr += s.drain()
        key = results_key(allr)
        assert len(key) == len(set((k[0], k[1]) for k in key))  # no duplicates
        if base is None:
            base = key
        assert key == base, f"P={P}"


FX = json.load(open(os.path.join(HERE, "golden", "search_counts.json")))


@pytest.mark.parametrize("grid", [25, 49])
@pytest.mark.parametrize("legacy", [False, True], ids=["current", "legacy_death"])
def test_reference_benchmark_sweep_counts(grid, legacy):
    """The reference's benchmark sweep (bench_grand_release.ps1:31-45: cube,
    FE, K = 2, --min-cs 2 --max-cs-range 3, --ticks 20, 11^3 filled) on the
    GPU equals the CPU oracle's counts; with the superseded death test it also
    reproduces the CSV's unviable count (bench_results.csv:6-9)."""
    print(f"sweep grid={grid} legacy={legacy}", flush=True)
    s = cv.Search(grid, cv.FACES | cv.EDGES, 2, 2, 3, 20, spawn=11, spawn_mode=cv.SPAWN_FILLED,
                  sim_flags=cv.SIM_LEGACY_DEATH if legacy else 0)
    s.run()
    p = s.progress()
    want = FX["oracle"][f"{grid}{'_legacy_death' if legacy else ''}"]
    assert (p["total"], p["viable"], p["possible"], p["unviable"]) == (
        want["total"], want["viable"], want["possible"], want["unviable"])
    if legacy and grid == 49:
        assert p["unviable"] == FX["csv"]["49"]["unviable"]


def test_cli_search_sweep(tmp_path):
    out = subprocess.run([CLI, "--mode", "search", "--shape", "cube", "--grid", "25", "--neighborhood", "FE",
                          "--chunks-per-config", "4", "--search-workgroup-size", "32", "--max-rule-bits", "2",
                          "--min-cs", "2", "--max-cs-range", "3", "--ticks", "20", "--dual-gpu", "--headless",
                          "--output", str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    # the regex of benchmarks/scripts/release/bench_grand_release.ps1:56
    rx = re.compile(r"viable: (\d+), possible: (\d+), unviable: (\d+).+_c(\d+)_.+_gpu(\d+)_viable")
    tot = np.zeros(3, np.int64)
    seen = set()
    for line in out.stdout.splitlines():
        m = rx.search(line)
        if m and (m.group(4), m.group(5)) not in seen:
            seen.add((m.group(4), m.group(5)))
            tot += [int(m.group(1)), int(m.group(2)), int(m.group(3))]
    want = FX["oracle"]["25"]
    assert len(seen) == 4
    assert list(tot) == [want["viable"], want["possible"], want["unviable"]]
    agg = tmp_path / "cube" / "cs2-3_FE_t20_b2_aggregate.jsonl"
    rows = [json.loads(l) for l in open(agg)]
    assert len(rows) == 71532  # every rule evaluated once at this single grid / seed
    assert sum(r["viableRuns"] for r in rows) == want["viable"]


def test_cli_jsonl_bytes(tmp_path):
    """The JSONL bytes of a sweep with 3 random-spawn seed runs per rule.

    Per-evaluation lines: the reference writes nlohmann::json's dump() of an
    object of unsigned integers (searchSystem.cpp:959-976; nlohmann_json
    v3.12.0, whose object is a std::map): compact, keys in byte order, decimal
    integers, one line each.  That is Python's compact sorted dump.

    Aggregate lines (ruleAggregator.h:104-131): fixed key order, the float
    fields as float32 quotients written by a default ostream, i.e. printf
    "%g".  The reference's map is unordered, so lines compare as a set.
    Each expected aggregate line is rebuilt here from the per-evaluation
    lines."""
    out = subprocess.run([CLI, "--mode", "search", "--shape", "cube", "--grid", "9", "--neighborhood", "FE",
                          "--chunks-per-config", "3", "--max-rule-bits", "1", "--min-cs", "2", "--max-cs-range", "3",
                          "--ticks", "12", "--spawn", "7", "--spawn-mode", "random", "--seed-runs", "3", "--headless",
                          "--output", str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    evals, agg_lines = [], None
    cats = {"_viable.jsonl": 2, "_possible.jsonl": 1, "_unviable.jsonl": 0}
    for root, _, files in os.walk(tmp_path):
        for f in files:
            raw = open(os.path.join(root, f), "rb").read().decode("ascii")
            if f.endswith("_aggregate.jsonl"):
                agg_lines = raw.splitlines(keepends=True)
                continue
            cat = next(c for sfx, c in cats.items() if f.endswith(sfx))
            for line in raw.splitlines(keepends=True):
                obj = json.loads(line)
                assert line == json.dumps(obj, sort_keys=True, separators=(",", ":")) + "\n", line
                assert all(type(v) is int and v >= 0 for v in obj.values())
                evals.append((obj, cat))
    assert evals and len(evals) % 3 == 0
    pm = (1 << 60) - 1
    stats = {}
    for obj, cat in evals:
        b, sv = obj["birthAndMaxCellStateRules"], obj["survivalAndNeighborhoodRules"]
        key = (b & pm, sv & pm, ((b >> 60) & 0xF) + 1, sv >> 60, 0)
        st = stats.setdefault(key, {"n": 0, 2: 0, 1: 0, 0: 0, "min": 2**32 - 1, "max": 0, "sum": 0})
        st["n"] += 1
        st[cat] += 1
        t = obj["ticksSurvived"]
        st["min"], st["max"], st["sum"] = min(st["min"], t), max(st["max"], t), st["sum"] + t
    assert len(stats) == len(evals) // 3 and all(st["n"] == 3 for st in stats.values())

    def g(num, den):
        return "%g" % float(np.float32(num) / np.float32(den))

    want = sorted(
        f'{{"birth":{k[0]},"survival":{k[1]},"maxCellState":{k[2]},"neighborhoodFlags":{k[3]},"shapeId":{k[4]},'
        f'"totalRuns":{st["n"]},"viableRuns":{st[2]},"possibleRuns":{st[1]},"unviableRuns":{st[0]},'
        f'"viabilityRate":{g(st[2], st["n"])},"minTicksSurvived":{st["min"]},"maxTicksSurvived":{st["max"]},'
        f'"avgTicksSurvived":{g(st["sum"], st["n"])}}}\n'
        for k, st in stats.items())
    assert agg_lines is not None and sorted(agg_lines) == want
    # fractions occur, so the float form is exercised
    assert any("." in l for l in want)


def test_grouped_chunks_equal_chunks_alone():
    """cave_search_step_group (chunks sharing launches, as the reference runs
    the chunks its VRAM budget admits side by side, searchMode.cpp:349-404)
    gives every chunk the results it gets alone."""
    def chunk(p, P):
        return cv.Search(13, cv.FACES | cv.EDGES, 2, 1, 3, 12, spawn=5, spawn_mode=cv.SPAWN_RANDOM,
                         density_index=6, seed_index=2, partition_index=p, partition_count=P)

    P = 7
    alone = []
    for p in range(P):
        c = chunk(p, P)
        c.run()
        alone.append((c.progress(), results_key(c.drain())))
        c.close()
    group = [chunk(p, P) for p in range(P)]
    cv.search_step_group(group)
    for p, c in enumerate(group):
        assert (c.progress(), results_key(c.drain())) == alone[p], f"partition {p}"
        assert not c.running
        c.close()


def cli_totals(stdout):
    rx = re.compile(r"viable: (\d+), possible: (\d+), unviable: (\d+).+_c(\d+)_.+_gpu(\d+)_viable")
    tot, seen = np.zeros(3, np.int64), set()
    for line in stdout.splitlines():
        m = rx.search(line)
        if m and (m.group(4), m.group(5)) not in seen:
            seen.add((m.group(4), m.group(5)))
            tot += [int(m.group(1)), int(m.group(2)), int(m.group(3))]
    return len(seen), list(tot)


def test_cli_vram_budget_groups(tmp_path):
    """--gpu-vram-budget-gb: at 25^3 a chunk of the benchmark sweep is
    estimated at 16.9 MB (permutationMath.h:144-153), so the default 40 GB
    admits all 64 chunks into one group and the 0.5 GB minimum 31 per group
    (3 groups).  Both give the oracle's counts and the same aggregate."""
    args = ["--mode", "search", "--grid", "25", "--neighborhood", "FE", "--max-rule-bits", "2", "--min-cs", "2",
            "--max-cs-range", "3", "--ticks", "20", "--headless"]
    runs = {}
    for label, budget in (("wide", "40"), ("narrow", "0.5")):
        out = subprocess.run([CLI, *args, "--gpu-vram-budget-gb", budget, "--output", str(tmp_path / label)],
                             capture_output=True, text=True, timeout=300)
        assert out.returncode == 0, out.stderr
        groups = [l for l in out.stdout.splitlines() if "Running chunks" in l]
        assert len(groups) == (1 if label == "wide" else 3), groups
        n, tot = cli_totals(out.stdout)
        want = FX["oracle"]["25"]
        assert n == 64 and tot == [want["viable"], want["possible"], want["unviable"]]
        agg = tmp_path / label / "cube" / "cs2-3_FE_t20_b2_aggregate.jsonl"
        runs[label] = sorted(open(agg).read().splitlines())
    assert runs["wide"] == runs["narrow"]
