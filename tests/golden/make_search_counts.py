"""Writes tests/golden/search_counts.json: the CPU oracle's classification
counts for the reference's benchmark sweep (cube, FE, --max-rule-bits 2,
--min-cs 2 --max-cs-range 3, --ticks 20, 11^3 filled spawn), at 25^3 and 49^3,
with the current death test and with the superseded one.

    python tests/golden/make_search_counts.py     # ~20 min on 8 cores

The reference's own numbers for the same sweep are in
benchmarks/reference/debug/bench_results.csv:2-9 (copied below as `csv`).
"""
import ctypes as C
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))
import oracle as O  # noqa: E402

CSV = {"25": {"viable": 848, "possible": 29032, "unviable": 41652, "total": 71532},
       "49": {"viable": 12570, "possible": 34230, "unviable": 24732, "total": 71532}}


def main():
    L = O.lib()
    L.orc_run_search_ex.restype = C.c_uint64
    L.orc_run_search_ex.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_double, C.c_uint32, C.c_int, C.POINTER(C.c_uint64), C.c_int]
    out = {"sweep": {"flags": 3, "max_rule_bits": 2, "min_cs": 2, "max_cs": 3, "ticks": 20, "spawn": 11,
                     "spawn_mode": "filled"}, "csv": CSV, "oracle": {}}
    for grid in (25, 49):
        for legacy in (0, 1):
            c = (C.c_uint64 * 3)()
            t = time.time()
            n = L.orc_run_search_ex(grid, 3, 2, 2, 3, 20, 11, 1, 0.5, 0, os.cpu_count(), c, legacy)
            key = f"{grid}{'_legacy_death' if legacy else ''}"
            out["oracle"][key] = {"total": int(n), "unviable": int(c[0]), "possible": int(c[1]), "viable": int(c[2])}
            print(key, out["oracle"][key], f"{time.time() - t:.0f}s", flush=True)
    with open(os.path.join(HERE, "search_counts.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
