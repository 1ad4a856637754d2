"""Benchmark of the CAVE MI355X engine.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu] [--no-search] [--no-serve]

A step is one tick of the single-grid workload: a 512^3 grid, FEV
neighbourhood (26 neighbours), C = 8 states (M = 7), random birth/survival
bitsets over counts 1..26 (p = 0.15 per bit, splitmix64 seed 2), uniform
Bernoulli(0.2) initial cells (cell-hash seed 42).  Inputs are resident in HBM
before the timed region.  `value` is cell-updates/s over the K timed ticks.

The same JSON line also carries the rule-search rate at 49^3 x 20 ticks
(4096 configs x 32 seeds, FE, C in 2..5), the uneven-size autotuned search
(2048 configs, edge 24..160) with its per-bucket achieved fractions, online
serving (Poisson 2000 requests/s for 60 s into the resident kernel: throughput,
p50 / p99 latency, oracle spot checks) and the host-CPU baselines, each timed
on a bounded sample with the fast CPU form (bit-sliced, threaded), which is
itself checked against the scalar oracle and the GPU results.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "cell-updates/sec + % HBM roofline; rule-configs/sec at 49^3×20 ticks; p50/p99 latency"
HBM_PEAK_GBS = 8000.0
GRID = 512


def log(msg: str):
    """Progress line on stderr (the JSON result line is the only stdout)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def cpu_info():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def workload_rule():
    import cave_amd as cv

    rng = cv.Rng(2)
    birth = rng.rule(26, 0.15)
    surv = rng.rule(26, 0.15)
    return birth, surv


def load_traffic(summary: str):
    """HBM bytes per launch of the dominant kernel, from the committed rocprof
    summary (FETCH_SIZE / WRITE_SIZE scaled by a plain copy of known size)."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":", 1)[1])
    return None


def run_step_bench(steps: int, warmup: int, dims=(GRID, GRID, GRID), wrap: bool = False):
    """Times `steps` ticks after `warmup` untimed ones.  Returns the wall
    seconds and device ms of the timed ticks, the kernel name, the cells after
    all warmup + steps ticks and the last tick's statistics (for the parity
    check against the CPU)."""
    import cave_amd as cv

    birth, surv = workload_rule()
    h = cv.Step(dims, 7, cv.FACES | cv.EDGES | cv.CORNERS | (cv.WRAP if wrap else 0), birth, surv,
                spawn=dims, spawn_mode=cv.SPAWN_RANDOM, density=0.2, seed=42)
    if warmup:
        h.run(warmup, stats=True)
    t0 = time.perf_counter()
    rows, dev_ms = h.run(steps, stats=True)
    wall = time.perf_counter() - t0
    kernel = "stream" if h.kernel == cv.KERNEL_STREAM else "gather"
    cells = h.cells()
    h.close()
    return wall, dev_ms, kernel, cells, (rows[-1].as_tuple() if rows else None)


def run_step_variant(dims, kernel_name: str, steps: int, warmup: int):
    """Event-timed launches of the step workload's rule on another grid or
    kernel form: the grid beyond the Infinity Cache (512 x 512 x 2048, a
    512 MiB ping-pong pair against a 256 MB cache), the nibble-layout form
    of the streaming kernel at 512^3, and the wrapping step (the W flag: every
    axis wraps, as in the render path's compute shader) at 512^3.  Returns the
    average launch time and its algorithmic HBM rate (1 B per cell-update)."""
    import cave_amd as cv

    birth, surv = workload_rule()
    kernel = {"stream": cv.KERNEL_STREAM, "stream_nibble": cv.KERNEL_STREAM_NIBBLE,
              "stream_wrap": cv.KERNEL_STREAM}[kernel_name]
    flags = cv.FACES | cv.EDGES | cv.CORNERS | (cv.WRAP if kernel_name == "stream_wrap" else 0)
    with cv.Step(dims, 7, flags, birth, surv, spawn=dims,
                 spawn_mode=cv.SPAWN_RANDOM, density=0.2, seed=42, kernel=kernel) as h:
        h.run(warmup, stats=False)
        t0 = time.perf_counter()
        _, dev_ms = h.run(steps, stats=True)
        wall = time.perf_counter() - t0
    cells = dims[0] * dims[1] * dims[2]
    launch_s = dev_ms / 1e3 / steps
    gbs = cells / launch_s / 1e9
    summary = {((GRID, GRID, 4 * GRID), "stream"): "step_stream_512x512x2048_FEV_C8.md",
               ((GRID, GRID, GRID), "stream_wrap"): "step_stream_wrap_512_FEV_C8.md",
               ((1024, 1024, 1024), "stream"): "step_stream_1024_FEV_C8.md"}.get((tuple(dims), kernel_name))
    return {"grid": list(dims), "kernel": f"step_{kernel_name}", "ticks": steps, "warmup": warmup,
            "traffic": load_traffic(summary) if summary else None,
            "avg_launch_us": launch_s * 1e6, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
            "cell_updates_per_s": cells * steps / wall, "state_bytes_ping_pong": cells}


def cpu_threads() -> int:
    """Host threads for the CPU baseline: the CPUs this process may run on,
    at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def oracle_module():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    return O


def fast_sims(specs):
    """The fast CPU form's simulation list (oracle SIM_NP records) for specs."""
    import numpy as np

    O = oracle_module()
    a = np.zeros(len(specs), dtype=O.SIM_NP)
    for k, f in enumerate(("X", "Y", "Z")):
        a[f] = specs["dims"][:, k]
    for k, f in enumerate(("sx", "sy", "sz")):
        a[f] = specs["spawn"][:, k]
    a["max_ticks"] = specs["ticks"]
    for f in ("M", "flags", "birth", "survival", "spawn_mode", "seed", "density"):
        a[f] = specs[f]
    return a


def cpu_step_baseline(gpu_cells=None, gpu_ticks: int = 0, gpu_last=None, target_s: float = 10.0):
    """Host-CPU baseline on the same workload: the fast CPU form (bit-sliced
    64-cell words, z slabs over threads) steps the 512^3 grid from the same
    initial cells for at least gpu_ticks ticks and about target_s seconds.
    After exactly gpu_ticks ticks its cells and last-tick statistics are
    compared with the GPU's (the parity record).  The fast form is itself
    checked against the scalar oracle on the first tick of a 256^3 grid with
    the same rule and initial density; that tick also times the scalar oracle
    (one thread)."""
    import numpy as np

    O = oracle_module()
    birth, surv = workload_rule()
    n = cpu_threads()
    E = GRID // 2
    small = O.init_cells(E, E, E, 7, (E, E, E), 0, 0.2, 42)
    t0 = time.perf_counter()
    ref, _ = O.step(small, E, E, E, (6 << 60) | birth, (7 << 60) | surv)
    scalar_s = time.perf_counter() - t0
    g = O.FastGrid(small, E, E, E)
    g.run(7, 7, birth, surv, 1, n_threads=n, stats=False)
    assert np.array_equal(g.cells(), ref), "fast CPU form / scalar oracle disagree"
    g.close()
    cells = O.init_cells(GRID, GRID, GRID, 7, (GRID, GRID, GRID), 0, 0.2, 42)
    g = O.FastGrid(cells, GRID, GRID, GRID)
    del cells
    parity = None
    ticks, t0 = 0, time.perf_counter()
    while True:
        if ticks < gpu_ticks:
            k = min(64, gpu_ticks - ticks)
            rows = g.run(7, 7, birth, surv, k, n_threads=n, stats=ticks + k == gpu_ticks)
            ticks += k
            if ticks == gpu_ticks and gpu_cells is not None:
                same_cells = bool(np.array_equal(g.cells(), gpu_cells))
                same_stats = gpu_last is None or rows[-1].as_tuple() == tuple(gpu_last)
                parity = {"ticks": gpu_ticks, "match": same_cells and same_stats, "cells_equal": same_cells,
                          "last_tick_stats_equal": same_stats,
                          "against": "fast CPU form, same initial cells and rule, after every warmup and "
                                     "timed tick"}
                if not parity["match"]:
                    log(f"PARITY FAILURE: GPU and CPU grids differ after {gpu_ticks} ticks")
        else:
            g.run(7, 7, birth, surv, 4, n_threads=n, stats=False)
            ticks += 4
        if ticks >= gpu_ticks and time.perf_counter() - t0 >= target_s:
            break
    dt = time.perf_counter() - t0
    g.close()
    base = {"value": GRID ** 3 * ticks / dt, "unit": "cell-updates/s", "cores": n, "kind": "port",
            "sample": f"{ticks} ticks of the 512^3 FEV C=8 workload, fast CPU form (bit-sliced 64-cell words, "
                      f"z slabs over {n} threads); checked equal to the scalar oracle on the first tick of a "
                      f"{E}^3 grid of the same rule and density",
            "scalar_oracle": {"value": E ** 3 / scalar_s, "unit": "cell-updates/s", "cores": 1,
                              "sample": f"1 tick of a {E}^3 grid of the same rule and density, scalar per-cell "
                                        "oracle"}}
    return base, parity


def active_region(specs, res, dev_ms, nominal, tm, label):
    """What the active region saved on a timed batch: the cells the ticks
    computed (each tick computes the cells within one cell of a cell live so
    far; CaveSimTiming.computed_cells) against the nominal X*Y*Z per tick, the
    HBM-equivalent fraction of the computed cells, and the same list rerun
    with CAVE_SIM_FULL_GRID (every cell every tick), whose results must be
    equal."""
    import numpy as np

    import cave_amd as cv

    computed = int(tm["computed_cells"].astype(np.int64).sum())
    full = np.array(specs, copy=True)
    full["sim_flags"] |= cv.SIM_FULL_GRID
    fres, fms = cv.batch_run_np(full, engine=cv.ENGINE_PERSISTENT)
    same = all(np.array_equal(fres[f], res[f]) for f in
               ("died", "born", "ticks_survived", "status", "max_alive", "final_alive", "classification",
                "ticks_run", "last"))
    if not same:
        log(f"{label}: full-grid and active-region results differ")
    peak = HBM_PEAK_GBS * 1e9
    return {"computed_cells": computed, "computed_fraction": computed / nominal,
            "computed_cell_updates_per_s": computed / (dev_ms / 1e3),
            "roofline_frac_computed": computed / (dev_ms / 1e3) / peak,
            "full_grid": {"device_ms": fms, "speedup_of_region": fms / dev_ms,
                          "roofline_frac": nominal / (fms / 1e3) / peak, "results_equal": same},
            "note": "a tick computes only the cells within one cell of a cell live so far (the others are 0 "
                    "and stay 0); value and cell_updates count every cell, X*Y*Z per tick, as the reference "
                    "computes them; computed_* count the cells computed"}


def run_search_bench(n_configs=4096, seeds=32, plain_configs=64, cpu=True, block_threads=0):
    """Rule-search rate at 49^3 x 20 ticks (FE, C in 2..5, <= 2 + 2 rule bits,
    32 seeds per config).  Persistent batched engine over the whole list; the
    plain one-launch-per-config engine and the CPU oracle on bounded samples."""
    import numpy as np

    import cave_amd as cv
    from cave_amd import workloads as W

    specs = W.search_uniform(n_configs, seeds)
    cells = 49 ** 3
    cv.batch_run_np(specs[: seeds * 8])  # warm-up (code objects, allocator)
    # The timed call is the second on the whole list: the first also grows
    # the engine's buffers to the list's size (its wall time is reported too).
    t0 = time.perf_counter()
    cv.batch_run_np(specs, engine=cv.ENGINE_PERSISTENT, block_threads=block_threads, timing=True)
    wall_first = time.perf_counter() - t0
    t0 = time.perf_counter()
    res, dev_ms, tm, threads = cv.batch_run_np(specs, engine=cv.ENGINE_PERSISTENT, block_threads=block_threads,
                                               timing=True)
    wall = time.perf_counter() - t0
    agg = W.aggregate_by_config(specs, res, seeds)
    # one aggregate JSONL record per rule config (RuleAggregator,
    # ruleAggregator.h:69-135): the seeds of a config fold into its record
    import tempfile

    t1 = time.perf_counter()
    ra = cv.Aggregate()
    ra.record_batch(specs, res)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "uniform_search_aggregate.jsonl")
        ra.write_jsonl(path)
        with open(path) as f:
            records = sum(1 for _ in f)
    jsonl_ms = (time.perf_counter() - t1) * 1e3
    ra.close()
    unique_rules = len({(int(b), int(s), int(m)) for b, s, m in
                        zip(specs["birth"][::seeds], specs["survival"][::seeds], specs["M"][::seeds])})
    updates = int(res["ticks_run"].astype(np.int64).sum()) * cells
    out = {
        "metric": "rule-configs/sec at 49^3x20 ticks",
        "value": n_configs / (dev_ms / 1e3),
        "value_wall": n_configs / wall,
        "unit": "configs/s",
        "configs": n_configs, "seeds_per_config": seeds, "sims": len(specs),
        "device_ms": dev_ms, "wall_ms": wall * 1e3, "wall_ms_first_call": wall_first * 1e3,
        "block_threads": threads,
        "cell_updates": updates, "cell_updates_per_s": updates / (dev_ms / 1e3),
        "early_exit": "reference status-1 early death on",
        "classes": {"viable": int(agg["viable"].sum()), "possible": int(agg["possible"].sum()),
                    "unviable": int(agg["unviable"].sum())},
        "aggregate_jsonl": {"records": records, "unique_rule_configs": unique_rules, "host_ms": jsonl_ms,
                            "note": "one RuleAggregator record per distinct rule (configs drawing the same "
                                    "rule fold together, as in the reference); written after the timed launch"},
    }
    out["active_region"] = active_region(specs, res, dev_ms, updates, tm, "search")
    log(f"  persistent engine {dev_ms:.2f} ms (every cell: {out['active_region']['full_grid']['device_ms']:.2f} ms); "
        f"plain engine on {plain_configs} configs")
    sub = specs[: plain_configs * seeds]
    pres, pms = cv.batch_run_np(sub, engine=cv.ENGINE_PLAIN)
    assert np.array_equal(pres["ticks_run"], res["ticks_run"][: len(sub)]), "plain / persistent disagree"
    out["plain"] = {"value": plain_configs / (pms / 1e3), "unit": "configs/s", "device_ms": pms,
                    "sample": f"first {plain_configs} configs x {seeds} seeds, one step + one manager "
                              "launch per sim per tick"}
    out["speedup_vs_plain"] = out["value"] / out["plain"]["value"]
    if cpu:
        O = oracle_module()
        n = cpu_threads()
        ticks, _, done, cpu_updates, dt = O.run_sims_fast(fast_sims(specs), n_threads=n, budget_s=10.0)
        ran = ticks != 0xFFFFFFFF
        assert np.array_equal(ticks[ran], res["ticks_run"][ran]), "fast CPU form / GPU disagree"
        out["cpu_baseline"] = {"value": done / seeds / dt, "unit": "configs/s",
                               "cell_updates_per_s": cpu_updates / dt, "cores": n, "kind": "port",
                               "sample": f"{done} sims ({done / seeds:.1f} configs) of the same list in list order, "
                                         f"fast CPU form, one sim per thread on {n} threads; ticks run equal "
                                         "the GPU's"}
    return out


def run_reference_sweep(partitions=(1, 64)):
    """The reference's own published workload, run as its release benchmark
    script runs it (benchmarks/scripts/release/bench_grand_release.ps1:30-66):
    `--mode search --shape cube --grid 49 --neighborhood FE
    --chunks-per-config P --search-workgroup-size 32 --max-rule-bits 2
    --min-cs 2 --max-cs-range 3 --ticks 20 --dual-gpu --headless`, one fresh
    process per P, timed as whole-process wall (start-up, device
    initialisation and output files included), the per-chunk counts summed
    from the log with the script's regex.  71,532 rule configs per sweep; the
    reference's published rates are 339 configs/s (Release, P = 1;
    bench_grand_results.csv:2) and 792 unique configs/s (Debug, P = 64;
    bench_grand_results_DEBUG.csv:8), both on 2x Quadro RTX 8000."""
    import re
    import shutil
    import subprocess
    import tempfile

    cli = os.path.join(ROOT, "cave_amd", "_lib", "cave")
    with open(os.path.join(ROOT, "tests", "golden", "search_counts.json")) as f:
        want = json.load(f)["oracle"]["49"]
    rx = re.compile(r"viable: (\d+), possible: (\d+), unviable: (\d+).+_c(\d+)_.+_gpu(\d+)_viable")
    published = {1: 339.0, 64: 792.0}
    out = {"workload": "cube 49^3, FE, --max-rule-bits 2, --min-cs 2 --max-cs-range 3, --ticks 20, 11^3 filled "
                       "spawn (CLI defaults), whole-process wall per sweep", "runs": []}
    for P in partitions:
        d = tempfile.mkdtemp(prefix="cave_sweep_")
        try:
            t0 = time.perf_counter()
            r = subprocess.run([cli, "--mode", "search", "--shape", "cube", "--grid", "49", "--neighborhood", "FE",
                                "--chunks-per-config", str(P), "--search-workgroup-size", "32", "--max-rule-bits",
                                "2", "--min-cs", "2", "--max-cs-range", "3", "--ticks", "20", "--dual-gpu",
                                "--headless", "--output", d], capture_output=True, text=True, timeout=600)
            wall = time.perf_counter() - t0
        finally:
            shutil.rmtree(d, ignore_errors=True)
        if r.returncode != 0:
            raise RuntimeError(f"cave --mode search failed ({r.returncode}): {r.stderr[-2000:]}")
        tot, seen = [0, 0, 0], set()
        for line in r.stdout.splitlines():
            m = rx.search(line)
            if m and (m.group(4), m.group(5)) not in seen:
                seen.add((m.group(4), m.group(5)))
                for k in range(3):
                    tot[k] += int(m.group(k + 1))
        n = sum(tot)
        match = tot == [want["viable"], want["possible"], want["unviable"]]
        if not match:
            log(f"SWEEP COUNT MISMATCH at P={P}: {tot} vs the oracle's {want}")
        out["runs"].append({"chunks_per_config": P, "chunks": len(seen), "rule_configs": n, "wall_s": wall,
                            "configs_per_s": n / wall, "reference_configs_per_s": published[P],
                            "vs_reference": n / wall / published[P],
                            "counts": {"viable": tot[0], "possible": tot[1], "unviable": tot[2]},
                            "counts_match_oracle": match})
    return out


def run_erd_search_bench(n_configs=4096, seeds=32, plain_configs=64):
    """The uniform rule search on the ERD cell shape (49^3, F + E = 12 face +
    4 edge neighbours, <= 2 + 2 rule bits over counts 1..16, C in 2..5, 32
    seeds, 20 ticks) on the persistent engine's bit-sliced ERD tick, beside
    the per-config engine (the reference-form ERD kernel, one step and one
    manager launch per simulation per tick) on the first configs."""
    import numpy as np

    import cave_amd as cv
    from cave_amd import workloads as W

    specs = W.search_uniform_erd(n_configs, seeds)
    cells = 49 ** 3
    cv.batch_run_np(specs[: seeds * 8])  # warm-up
    res, dev_ms, tm, _ = cv.batch_run_np(specs, engine=cv.ENGINE_PERSISTENT, timing=True)
    updates = int(res["ticks_run"].astype(np.int64).sum()) * cells
    region = active_region(specs, res, dev_ms, updates, tm, "search_erd")
    sub = specs[: plain_configs * seeds]
    pres, pms = cv.batch_run_np(sub, engine=cv.ENGINE_PLAIN)
    same = bool(np.array_equal(pres["ticks_run"], res["ticks_run"][: len(sub)]) and
                np.array_equal(pres["final_alive"], res["final_alive"][: len(sub)]))
    if not same:
        log("ERD persistent / per-config engines disagree")
    return {"metric": "rule-configs/sec at 49^3x20 ticks, ERD cell shape",
            "value": n_configs / (dev_ms / 1e3), "unit": "configs/s", "configs": n_configs,
            "seeds_per_config": seeds, "device_ms": dev_ms, "cell_updates_per_s": updates / (dev_ms / 1e3),
            "hbm_equivalent_frac": updates / (dev_ms / 1e3) / (HBM_PEAK_GBS * 1e9),
            "active_region": region,
            "plain": {"value": plain_configs / (pms / 1e3), "unit": "configs/s", "device_ms": pms,
                      "sample": f"first {plain_configs} configs x {seeds} seeds, the reference-form ERD kernel: "
                                "one step + one manager launch per sim per tick",
                      "results_equal": same},
            "speedup_vs_plain": (n_configs / (dev_ms / 1e3)) / (plain_configs / (pms / 1e3))}


EDGE_BUCKETS = ["24-31", "32-47", "48-63", "64-95", "96-127", "128-160"]
FLAG_NAMES = {1: "F", 3: "FE", 7: "FEV"}


def run_uneven_bench(n_configs=2048, cpu=True):
    """Batched rule search with uneven sizes: 2048 configs, edge log-uniform
    in [24, 160], F : FE : FEV = 1 : 2 : 1, C in 2..16, bits p = 0.1, ticks in
    [20, 200], central cube of edge/2 Bernoulli(0.25).  The autotuner runs in
    the untimed first call; the timed call reports per-(edge bucket,
    neighbourhood) block time and achieved fraction of the block's share of
    the HBM roofline (1 B per cell-update)."""
    import numpy as np

    import cave_amd as cv
    from cave_amd import workloads as W

    specs = W.search_uneven(n_configs)
    cells = specs["dims"].astype(np.int64).prod(1)
    t0 = time.perf_counter()
    cv.batch_run_np(specs)  # autotuner calibration + warm-up
    tune_s = time.perf_counter() - t0
    log(f"  autotuner + warm-up call {tune_s:.2f} s")
    t0 = time.perf_counter()
    res, dev_ms, tm, threads = cv.batch_run_np(specs, engine=cv.ENGINE_PERSISTENT, timing=True)
    wall = time.perf_counter() - t0
    _, cus, _ = cv.device_info()
    resident = cus * max(1, 1024 // threads)
    block_share = HBM_PEAK_GBS * 1e9 / resident  # B/s of HBM per resident block
    ups = res["ticks_run"].astype(np.int64) * cells
    edge = np.cbrt(cells.astype(np.float64))
    bidx = np.digitize(edge, [32, 48, 64, 96, 128])
    block_s = (tm["end_us"] - tm["start_us"]) * tm["blocks"] / 1e6
    comp = tm["computed_cells"].astype(np.int64)
    buckets = []
    for b in range(6):
        for fl in (1, 3, 7):
            m = (bidx == b) & (specs["flags"] == fl)
            if not m.any():
                continue
            rate = ups[m].sum() / block_s[m].sum()
            buckets.append({"edge": EDGE_BUCKETS[b], "neighborhood": FLAG_NAMES[fl], "configs": int(m.sum()),
                            "cell_updates": int(ups[m].sum()), "block_ms": float(block_s[m].sum() * 1e3),
                            "split_configs": int((tm["blocks"][m] > 1).sum()),
                            "items_per_thread": int(np.bincount(tm["items_per_thread"][m]).argmax()),
                            "cell_updates_per_block_s": rate, "achieved_frac": rate * 1.0 / block_share,
                            "computed_fraction": float(comp[m].sum() / ups[m].sum()),
                            "achieved_frac_computed": comp[m].sum() / block_s[m].sum() / block_share})
    out = {
        "metric": "rule-configs/sec, uneven sizes (edge 24..160, F/FE/FEV, C 2..16, 20..200 ticks)",
        "value": n_configs / (dev_ms / 1e3), "value_wall": n_configs / wall, "unit": "configs/s",
        "configs": n_configs, "device_ms": dev_ms, "cell_updates": int(ups.sum()),
        "cell_updates_per_s": ups.sum() / (dev_ms / 1e3),
        "roofline_frac": ups.sum() * 1.0 / (dev_ms / 1e3) / (HBM_PEAK_GBS * 1e9),
        "block_threads": int(threads), "tuning_s": tune_s,
        "split_configs": int((tm["blocks"] > 1).sum()),
        "last_finish_ms": float(tm["end_us"].max() / 1e3), "p50_finish_ms": float(np.percentile(tm["end_us"], 50) / 1e3),
        "buckets": buckets,
    }
    out["active_region"] = active_region(specs, res, dev_ms, int(ups.sum()), tm, "search_uneven")
    log(f"  timed call {dev_ms:.2f} ms (every cell: {out['active_region']['full_grid']['device_ms']:.2f} ms)")
    if cpu:
        O = oracle_module()
        n = cpu_threads()
        ticks, _, done, cpu_ups, dt = O.run_sims_fast(fast_sims(specs), n_threads=n, budget_s=10.0)
        ran = ticks != 0xFFFFFFFF
        assert np.array_equal(ticks[ran], res["ticks_run"][ran]), "fast CPU form / GPU disagree"
        cpu_rate = cpu_ups / dt
        out["cpu_baseline"] = {"value": cpu_rate / (ups.sum() / n_configs), "unit": "configs/s",
                               "cell_updates_per_s": cpu_rate, "cores": n, "kind": "port",
                               "sample": f"{done} simulations of the list in list order, fast CPU form, one "
                                         f"simulation per thread on {n} threads; ticks run equal the GPU's; "
                                         "configs/s = CPU cell-updates/s / mean executed cell-updates per config"}
    return out


def run_serving_bench(rate_hz=2000.0, duration_s=60.0, cpu=True, split_us=0, max_parts=0):
    """Online serving: Poisson arrivals at rate_hz for duration_s, each request
    a grid of edge 32..192 (log-uniform), F/FE/FEV, C in 2..16, bits p = 0.12,
    20..500 ticks, admitted against a 200 GB HBM budget into the resident
    persistent kernel.  Latency = completion seen by the host - submission.
    1% of the requests (every 100th) carry a final-cell hash; each one is
    rerun on the fast CPU form and compared."""
    import numpy as np

    import cave_amd as cv
    from cave_amd import workloads as W

    times, specs = W.serving_requests(rate_hz, duration_s)
    n = len(specs)
    sample = np.arange(0, n, 100)
    specs["sim_flags"][sample] |= cv.SIM_FINAL_HASH
    specs = np.ascontiguousarray(specs)
    srv = cv.Server(hbm_budget_bytes=200 << 30, split_target_us=split_us, max_parts=max_parts)
    for i in range(8):  # warm-up requests (code objects, first arena chunk)
        srv.submit_row(specs, i, 1 << 40 | i)
    srv.drain()
    log(f"  serving {n} requests over {duration_s:.0f} s")
    done = {}
    t0 = time.perf_counter()
    nxt, last_log = 0, 0.0
    while len(done) < n:
        now = time.perf_counter() - t0
        while nxt < n and times[nxt] <= now:
            srv.submit_row(specs, nxt, nxt)
            nxt += 1
        for c in srv.poll():
            done[c.request_id] = c
        if now - last_log >= 10.0:
            last_log = now
            log(f"  t={now:5.1f} s submitted {nxt} finished {len(done)}")
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    cells = specs["dims"].astype(np.int64).prod(1)
    ticks_run = np.array([done[i].result.ticks_run for i in range(n)], dtype=np.int64)
    lat = np.array([done[i].finish_s - done[i].submit_s for i in range(n)]) * 1e3
    queue = np.array([done[i].admit_s - done[i].submit_s for i in range(n)]) * 1e3
    run = np.array([done[i].run_ms for i in range(n)])
    blocks = np.array([done[i].blocks for i in range(n)])
    bad = sum(done[i].status != 0 for i in range(n))
    ups = int((ticks_run * cells).sum())
    out = {
        "metric": "served rule evaluations: throughput and p50/p99 latency",
        "value": n / wall, "unit": "requests/s", "offered_rate": rate_hz, "requests": n, "wall_s": wall,
        "cell_updates_per_s": ups / wall,
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                       "max": float(lat.max()), "mean": float(lat.mean())},
        "admission_wait_ms": {"p50": float(np.percentile(queue, 50)), "p99": float(np.percentile(queue, 99))},
        "device_run_ms": {"p50": float(np.percentile(run, 50)), "p99": float(np.percentile(run, 99))},
        "split_requests": int((blocks > 1).sum()), "failed_requests": int(bad),
        "peak_resident_bytes": st["peak_resident_bytes"], "hbm_budget_bytes": 200 << 30,
    }
    if cpu:
        O = oracle_module()
        workers = cpu_threads()
        # every sampled request, costliest first so the threads finish together
        order = sample[np.argsort(-(ticks_run[sample] * cells[sample]), kind="stable")]
        t1 = time.perf_counter()
        cticks, cinfo, ran, cpu_ups, _, chash = O.run_sims_fast(fast_sims(specs[order]), n_threads=workers,
                                                              hashes=True)
        dt = time.perf_counter() - t1
        mism = []
        for k, i in enumerate(order):
            r = done[int(i)].result
            same = (cticks[k] == r.ticks_run and cinfo[k].finalAliveCount == r.info.finalAliveCount and
                    cinfo[k].maxAliveCount == r.info.maxAliveCount and
                    cinfo[k].totalNumberOfCellsThatDied == r.info.totalNumberOfCellsThatDied and
                    cinfo[k].totalNumberOfCellsThatWereBorn == r.info.totalNumberOfCellsThatWereBorn and
                    cinfo[k].status == r.info.status and int(chash[k]) == r.final_hash)
            if not same:
                mism.append(int(i))
        out["spot_checks"] = {"sampled": int(len(sample)), "checked": int(ran), "mismatches": len(mism),
                              "split_requests_checked": int((blocks[order] > 1).sum()),
                              "largest_checked_cell_updates": int((ticks_run[order] * cells[order]).max()),
                              "method": "every 100th request carries a final-cell hash; every one of them is rerun "
                                        f"on the fast CPU form (one request per thread, {workers} threads) and its "
                                        "ticks run, grid info and final-cell hash are compared",
                              "cpu_cell_updates_per_s": cpu_ups / dt, "cpu_seconds": dt, "cores": workers}
        if mism:
            log(f"SPOT-CHECK FAILURE: requests {mism[:10]} differ from the CPU")
        assert not mism, f"{len(mism)} served results differ from the CPU"
    return out


def device_copy_rate(nbytes: int, copies: int = 100, warm: int = 20):
    """Plain device-to-device copies of `nbytes` (hipMemcpyAsync on the null
    stream, HIP events around `copies` of them after `warm` untimed ones):
    the rate a kernel that reads and writes the same bytes once can expect.
    Returns (us per copy, GB/s read + written)."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so.7")  # the runtime the engine's library already loaded

    def ok(rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    src, dst, e0, e1 = C.c_void_p(), C.c_void_p(), C.c_void_p(), C.c_void_p()
    ok(hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)), "hipMalloc")
    try:
        ok(hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)), "hipMalloc")
        try:
            ok(hip.hipMemset(src, 0x5A, C.c_size_t(nbytes)), "hipMemset")
            ok(hip.hipEventCreate(C.byref(e0)), "hipEventCreate")
            ok(hip.hipEventCreate(C.byref(e1)), "hipEventCreate")
            copy = lambda: ok(hip.hipMemcpyAsync(dst, src, C.c_size_t(nbytes), 3, None), "hipMemcpyAsync")
            for _ in range(warm):
                copy()
            ok(hip.hipEventRecord(e0, None), "hipEventRecord")
            for _ in range(copies):
                copy()
            ok(hip.hipEventRecord(e1, None), "hipEventRecord")
            ok(hip.hipEventSynchronize(e1), "hipEventSynchronize")
            ms = C.c_float()
            ok(hip.hipEventElapsedTime(C.byref(ms), e0, e1), "hipEventElapsedTime")
            hip.hipEventDestroy(e0)
            hip.hipEventDestroy(e1)
        finally:
            hip.hipFree(dst)
    finally:
        hip.hipFree(src)
    us = ms.value * 1e3 / copies
    return us, 2 * nbytes / us / 1e3


def copy_calibration(mib: int, copies: int):
    """K plain device-to-device copies of a MIB-MiB buffer: each reads and
    writes exactly MIB MiB, the yardstick for the HBM byte counters.  Also
    times them (HIP events on torch's stream, after 5 untimed copies): the
    best a kernel moving the same bytes can expect."""
    import torch

    n = mib << 20
    src = torch.randint(0, 255, (n,), dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        dst.copy_(src)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(copies):
        dst.copy_(src)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) * 1e3 / max(1, copies)
    print(json.dumps({"copy_calibration": {"bytes_read_per_copy": n, "bytes_written_per_copy": n,
                                           "copies": copies, "us_per_copy": us,
                                           "tb_per_s_read_plus_write": 2 * n / us / 1e6}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-search", action="store_true")
    ap.add_argument("--search-configs", type=int, default=4096)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--step-dims", default=f"{GRID},{GRID},{GRID}",
                    help="grid of the single-grid step (X,Y,Z); other than 512^3 only with --no-cpu, "
                         "for profiling one size on its own")
    ap.add_argument("--step-wrap", action="store_true",
                    help="the single-grid step with the W flag (every axis wraps); only with --no-cpu, "
                         "for profiling the wrapping step on its own")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the 512x512x2048 and nibble-layout step timings")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-reference-sweep", action="store_true")
    ap.add_argument("--no-erd", action="store_true")
    ap.add_argument("--plain-configs", type=int, default=64)
    ap.add_argument("--search-threads", type=int, default=0,
                    help="block size of the uniform search launch (0: the tuner's choice); for profiling")
    ap.add_argument("--serve-split-us", type=int, default=0,
                    help="split target of the server (0 = its default)")
    ap.add_argument("--serve-max-parts", type=int, default=0,
                    help="most blocks per served request (0 = the server's default)")
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run K device-to-device copies of MIB MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE counters against a known byte count) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py runs on one GPU only: the whole grid lives on one MI355X and "
              "search/serving configs are independent (no multi-GPU path is built).", file=sys.stderr)
        sys.exit(2)

    if a.copy_calibration:
        copy_calibration(a.copy_calibration, a.steps)
        return

    import cave_amd as cv

    dev, cus, mem = cv.device_info()
    dims = tuple(int(v) for v in a.step_dims.split(","))
    if dims != (GRID, GRID, GRID) and not a.no_cpu:
        ap.error("--step-dims other than 512,512,512 needs --no-cpu (the CPU parity check runs the 512^3 grid)")
    if a.step_wrap and not a.no_cpu:
        ap.error("--step-wrap needs --no-cpu (the CPU parity check runs the non-wrapping workload)")
    variants = {}
    if not a.no_variants:
        # The 512^3 nibble-layout leg runs first: its 2000 warm-up ticks and
        # 500 timed ones also bring the GPU's clocks to their steady state for
        # an in-cache 512^3 step.  On an idle GPU the 20 ticks right after 5
        # warm-up ticks run ~7% slower while the clocks ramp up; right after
        # the memory-bound large-grid legs the in-cache step also runs slower
        # for a while (20.9 vs 20.0 us per launch), so those legs come after
        # the headline.
        log("step variant: nibble layout at 512^3")
        variants["step_nibble"] = run_step_variant((GRID, GRID, GRID), "stream_nibble", 500, 2000)
    log(f"single-grid step: {a.steps} ticks of {dims[0]}x{dims[1]}x{dims[2]} FEV")
    wall, dev_ms, kernel, gpu_cells, gpu_last = run_step_bench(a.steps, a.warmup, dims, a.step_wrap)
    cells = dims[0] * dims[1] * dims[2]
    value = cells * a.steps / wall
    launch_s = dev_ms / 1e3 / a.steps
    achieved = cells * 1.0 / launch_s / 1e9  # 0.5 B read + 0.5 B write per cell-update
    traffic = load_traffic("step_stream_512_FEV_C8.md") if kernel == "stream" and dims == (GRID,) * 3 else None
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall * 1e3 / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u4 (packed nibbles, integer)",
        "data": "synthetic (seeded cell hash + splitmix64 rules)",
        "config": {"workload": f"single grid {dims[0]}x{dims[1]}x{dims[2]}, FEV, C=8, random B/S p=0.15, "
                               "Bernoulli(0.2) init",
                   "grid": list(dims), "neighborhood": "FEV", "cell_states": 8,
                   "kernel": kernel, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"step_{kernel}", "avg_launch_us": launch_s * 1e6},
        "device": {"name": dev, "cus": cus, "hbm_gib": round(mem / 2**30, 1)},
    }
    # the yardstick beside the peak: a plain copy of one state buffer (cells / 2
    # bytes read and as many written, the bytes of one tick) on the same GPU
    copy_us, copy_gbs = device_copy_rate(cells // 2)
    out["roofline"].update({"copy_gbs": copy_gbs, "copy_us": copy_us, "frac_of_copy": achieved / copy_gbs,
                            "copy": f"hipMemcpyAsync of {cells // 2 >> 20} MiB device to device, 100 timed"})
    if not a.no_variants:
        log("step variant: the wrapping step (W) at 512^3")
        variants["step_wrap"] = run_step_variant((GRID, GRID, GRID), "stream_wrap", 500, 200)
        log("step variants: 512x512x2048 (beyond the Infinity Cache), large grids")
        variants["step_beyond_cache"] = run_step_variant((GRID, GRID, 4 * GRID), "stream", 200, 200)
        # the reference's large-grid test sizes (src/testing/testRunner.cpp:200-229) and 1024^3
        variants["step_large_grids"] = [run_step_variant((e, e, e), "stream", 50, 50) for e in (950, 1024, 1500, 2023)]
    out.update(variants)
    model, ncpu = cpu_info()
    if not a.no_cpu:
        log("host-CPU baseline on the single-grid workload")
        cb, parity = cpu_step_baseline(gpu_cells, a.warmup + a.steps, gpu_last)
        cb["host"] = f"{model} ({ncpu} logical CPUs visible)"
        out["cpu_baseline"] = cb
        out["parity"] = parity
    del gpu_cells
    if not a.no_search:
        log(f"rule search: {a.search_configs} configs x 32 seeds at 49^3")
        out["search"] = run_search_bench(a.search_configs, plain_configs=a.plain_configs, cpu=not a.no_cpu,
                                         block_threads=a.search_threads)
        if not a.no_uneven:
            log("rule search, uneven sizes: 2048 configs")
            out["search_uneven"] = run_uneven_bench(cpu=not a.no_cpu)
        if not a.no_erd:
            log("rule search on the ERD cell shape: 4096 configs x 32 seeds at 49^3")
            out["search_erd"] = run_erd_search_bench(plain_configs=a.plain_configs)
        if not a.no_reference_sweep:
            log("the reference's published sweep (49^3 FE, K = 2, cs 2-3, 20 ticks) through the CLI")
            out["reference_sweep"] = run_reference_sweep()
    if not a.no_serve:
        log(f"serving: Poisson {a.serve_rate:.0f} req/s for {a.serve_seconds:.0f} s")
        out["serving"] = run_serving_bench(a.serve_rate, a.serve_seconds, cpu=not a.no_cpu,
                                           split_us=a.serve_split_us, max_parts=a.serve_max_parts)
    log("done")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
