"""Writes the committed golden fixtures of the step path from the CPU oracle.

    python tests/golden/make_golden.py

Fixtures (JSON, small):
  step_49_FE_C2.json  -- 49^3, FE, C = 2 (M = 1), B = {5,6}, S = {4,5,6}, central
                         16^3 Bernoulli(0.35) spawn, seed 1234, 20 ticks: per-tick
                         counters, live, changed, bbox and canonical FNV-1a hash.
  step_small.json     -- 8^3..13^3 grids x {F, FE, FEV} x M in {1, 4, 15}: per-tick
                         counters and hashes for 6 ticks.
Each fixture is regenerated from the oracle, whose own pins live in
tests/test_oracle.py; the fixtures freeze the oracle's output so that a later
change of either side shows up as a diff.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(HERE)), "oracle"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402


def run(X, Y, Z, M, flags, birth, surv, spawn, mode, density, seed, ticks):
    cells = O.init_cells(X, Y, Z, M, spawn, mode, density, seed)
    w0 = ((M - 1) & 0xF) << 60 | birth
    w1 = flags << 60 | surv
    rows = [{"tick": 0, "hash": O.canonical_hash(cells, X, Y, Z)}]
    for t in range(ticks):
        cells, st = O.step(cells, X, Y, Z, w0, w1)
        d, dy, b, s, live, ch, bmin, bmax = st.as_tuple()
        rows.append({"tick": t + 1, "dead": d, "dying": dy, "born": b, "surviving": s, "live": live,
                     "changed": ch, "bbox_min": list(bmin), "bbox_max": list(bmax),
                     "hash": O.canonical_hash(cells, X, Y, Z)})
    return rows


def spec(X, Y, Z, M, flags, birth, surv, spawn, mode, density, seed, ticks):
    return {"dims": [X, Y, Z], "M": M, "flags": flags, "birth": birth, "survival": surv,
            "spawn": list(spawn), "spawn_mode": mode, "density": density, "seed": seed,
            "ticks": ticks}


def main():
    B = O.rule_string_to_bitmask("5,6")
    S = O.rule_string_to_bitmask("4,5,6")
    fe49 = spec(49, 49, 49, 1, 3, B, S, (16, 16, 16), 0, 0.35, 1234, 20)
    out = {"spec": fe49, "rows": run(49, 49, 49, 1, 3, B, S, (16, 16, 16), 0, 0.35, 1234, 20)}
    with open(os.path.join(HERE, "step_49_FE_C2.json"), "w") as f:
        json.dump(out, f, indent=0)

    rng = np.random.default_rng(20261203)
    cases = []
    for dims in [(8, 8, 8), (13, 9, 11), (9, 13, 8), (12, 12, 12)]:
        for flags in (1, 3, 7):
            for M in (1, 4, 15):
                n = O.lib().orc_neighbor_count(flags)
                birth = int(rng.integers(1, 1 << n))
                surv = int(rng.integers(1, 1 << n))
                sp = spec(*dims, M, flags, birth, surv, dims, 0, 0.4, int(rng.integers(0, 1 << 31)), 6)
                rows = run(*dims, M, flags, birth, surv, dims, 0, 0.4, sp["seed"], 6)
                cases.append({"spec": sp, "rows": rows})
    with open(os.path.join(HERE, "step_small.json"), "w") as f:
        json.dump(cases, f, indent=0)
    print("wrote", len(cases) + 1, "fixtures")


if __name__ == "__main__":
    main()
