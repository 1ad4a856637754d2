"""Tests candidate histories of the reference's benchmark CSV on the CPU.

benchmarks/reference/debug/bench_results.csv:2-9 gives, for one sweep (cube,
FE, --max-rule-bits 2, --min-cs 2 --max-cs-range 3, --ticks 20, 11^3 filled
spawn), the viable / possible / unviable counts at grids 25 and 49.  The
oracle (the current shaders) matches the 49^3 unviable count exactly under
the superseded death test but matches neither 25^3 count.  This script runs
the whole sweep (71,532 rules per grid) under each variant of
csv_provenance.c, classifies every rule under both death tests and, for
variant 5, with the tick budget checked before the death test, and writes the
counts next to the CSV's to csv_provenance.json.

    python tests/golden/csv_provenance.py             # full sweep, ~10 min on 8 cores
    python tests/golden/csv_provenance.py --sample 3000  # a seeded sample of rules
"""
import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import oracle as O  # noqa: E402

CSV = {25: {"viable": 848, "possible": 29032, "unviable": 41652},
       49: {"viable": 12570, "possible": 34230, "unviable": 24732}}
VARIANTS = {
    0: "current semantics (padded rows everywhere)",
    1: "own cell padded, neighbours read at the linear index (searchSimulationCube.slang:83-87)",
    2: "initial cells written linearly, shader padded (searchSystem.cpp:133-139)",
    3: "initial cells padded, prior 1D linear shader (searchSimulationCube.slang:7-12)",
    4: "prior 1D linear shader, neighbours bounds-checked on the linear index only",
    5: "current semantics, tick budget checked before the death test",
}


class Job(C.Structure):
    _fields_ = [("M", C.c_uint32), ("flags", C.c_uint32), ("birth", C.c_uint64), ("survival", C.c_uint64)]


def build_lib(tmp):
    src = os.path.join(HERE, "csv_provenance.c")
    so = os.path.join(tmp, "libcsv_provenance.so")
    subprocess.check_call(["gcc", "-O3", "-march=native", "-fPIC", "-shared", "-pthread", "-o", so, src])
    lib = C.CDLL(so)
    lib.cp_run.argtypes = [C.c_int, C.c_uint32, C.POINTER(Job), C.c_size_t, C.c_uint32, C.c_uint32, C.c_int,
                           C.POINTER(C.c_uint32)]
    return lib


def sweep_jobs():
    """The sweep's rules in the reference's order: maxCellState M = 2..3
    (--min-cs / --max-cs-range, searchMode.cpp:82), neighbourhood
    subsets F, E, FE of FE (searchMode.cpp:49-102), survival outer, birth
    inner over the K <= 2 bitmasks (permutationMath.h:56-81)."""
    jobs = []
    for M in (2, 3):
        for flags in (1, 2, 3):
            n = {1: 6, 2: 12, 3: 18}[flags]
            vb = O.enumerate_valid_bitmasks(n, 2)
            for s in vb:
                for b in vb:
                    jobs.append((M, flags, int(b), int(s)))
    return jobs


def classify(death_tick, M, ticks=20, budget_first=False):
    """searchSystem.cpp:695-711: 2 viable (status 2), 1 possible (died after
    more than M+1 ticks), 0 unviable."""
    if death_tick == 0 or (budget_first and death_tick >= ticks):
        return 2
    return 1 if death_tick > M + 1 else 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sample", type=int, default=0, help="seeded sample of rules (0: the whole sweep)")
    ap.add_argument("--variants", default="0,1,2,3,4")
    ap.add_argument("--grids", default="25,49")
    ap.add_argument("--threads", type=int, default=os.cpu_count())
    ap.add_argument("--dump", default="", help="also save every rule's death ticks to this .npz")
    a = ap.parse_args()
    jobs = sweep_jobs()
    assert len(jobs) == 71532, len(jobs)
    idx = np.arange(len(jobs))
    if a.sample:
        idx = np.sort(np.random.default_rng(7).choice(len(jobs), a.sample, replace=False))
    arr = (Job * len(idx))(*[Job(*jobs[i]) for i in idx])
    path = os.path.join(HERE, "csv_provenance.json")
    out = {"sweep": "cube, FE, K=2, maxCellState 2-3, 20 ticks, 11^3 filled spawn", "csv": {str(k): v for k, v in CSV.items()},
           "variants": {}}
    if os.path.exists(path):  # results accumulate over runs: each entry says how many rules it covers
        with open(path) as f:
            out["variants"] = json.load(f).get("variants", {})
    dumps = {}
    Ms = np.array([jobs[i][0] for i in idx])
    with tempfile.TemporaryDirectory() as tmp:
        lib = build_lib(tmp)
        for v in [int(x) for x in a.variants.split(",")]:
            for G in [int(x) for x in a.grids.split(",")]:
                res = (C.c_uint32 * (2 * len(idx)))()
                t = time.time()
                lib.cp_run(v, G, arr, len(idx), 20, 11, a.threads, res)
                r = np.frombuffer(res, dtype=np.uint32).reshape(-1, 2)
                dumps[f"v{v}_g{G}"] = r.copy()
                for vv, budget_first in ((v, False), (5, True)) if v == 0 else ((v, False),):
                    entry = {"rules": len(idx), "sample_seed": 7 if a.sample else None}
                    for test, col in (("current_death", 0), ("legacy_death", 1)):
                        cls = np.array([classify(int(d), int(m), budget_first=budget_first)
                                        for d, m in zip(r[:, col], Ms)])
                        entry[test] = {"viable": int((cls == 2).sum()), "possible": int((cls == 1).sum()),
                                       "unviable": int((cls == 0).sum())}
                        scale = len(idx) / len(jobs)
                        print(f"variant {vv} G={G} {test:14s} {entry[test]}  "
                              f"csv~{ {k: round(x * scale) for k, x in CSV[G].items()} }  {time.time() - t:.0f}s",
                              flush=True)
                    if vv == 0:
                        # unviable counts under every threshold "died within M + k ticks"
                        # (searchSystem.cpp:695-711 has k = 1): no k gives the 25^3 row's 41,652
                        entry["unviable_if_died_within_M_plus_k"] = {
                            test: {str(k): int(((r[:, col] > 0) & (r[:, col].astype(np.int64) <= Ms + k)).sum())
                                   for k in range(-3, 8)}
                            for test, col in (("current_death", 0), ("legacy_death", 1))}
                    var = out["variants"].setdefault(f"v{vv}", {})
                    var["what"] = VARIANTS[vv]
                    if not (a.sample and var.get(str(G), {}).get("rules", 0) > len(idx)):
                        var[str(G)] = entry  # a sample never replaces a whole-sweep entry
    if a.dump:
        np.savez(a.dump, jobs=np.array([jobs[i] for i in idx], dtype=np.uint64), **dumps)
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
