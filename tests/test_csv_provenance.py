"""The reference's benchmark CSV against candidate histories of its build.

benchmarks/reference/debug/bench_results.csv:2-9 reports viable / possible /
unviable counts of one sweep at grids 25 and 49.  tests/golden/
csv_provenance.py simulates that sweep under named semantics (the current
shaders and four earlier or defective states of them) and writes the counts
to tests/golden/csv_provenance.json.  These tests pin what that record
establishes:
  * the record's current-semantics counts are the oracle's
    (tests/golden/search_counts.json), so the harness restates the oracle;
  * the CSV's 49^3 unviable count is the current semantics' count under the
    superseded death test, while no recorded variant reproduces all six
    numbers;
  * the unviable counts of any local semantics cannot depend on the grid
    between 19^3 and 49^3 (a rule is unviable when it dies within M + 1 <= 4
    ticks, and an 11^3 spawn centred in a grid of edge >= 19 is at least 4
    cells from every face), checked rule by rule on a sample at 19^3 and
    25^3; the CSV's two unviable counts differ (41,652 vs 24,732), so its 25^3
    row cannot come from the semantics that give its 49^3 row.
"""
import ctypes as C
import json
import os
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
sys.path.insert(0, GOLDEN)


def _load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def test_current_semantics_match_the_oracle_counts():
    prov = _load("csv_provenance.json")["variants"]["v0"]
    oracle = _load("search_counts.json")["oracle"]
    for grid in ("25", "49"):
        assert prov[grid]["rules"] == 71532
        for test, key in (("current_death", grid), ("legacy_death", f"{grid}_legacy_death")):
            want = {k: oracle[key][k] for k in ("viable", "possible", "unviable")}
            assert prov[grid][test] == want, (grid, test)


def test_csv_49_unviable_is_the_superseded_death_test_and_no_variant_fits_all_six():
    rec = _load("csv_provenance.json")
    csv = rec["csv"]
    assert rec["variants"]["v0"]["49"]["legacy_death"]["unviable"] == csv["49"]["unviable"] == 24732
    for name, var in rec["variants"].items():
        for test in ("current_death", "legacy_death"):
            rows = [var[g][test] for g in ("25", "49") if g in var]
            full = all(var[g]["rules"] == 71532 for g in ("25", "49") if g in var)
            if len(rows) == 2 and full:
                assert not (rows[0] == csv["25"] and rows[1] == csv["49"]), (name, test)
            # the 25^3 row is fitted by no variant, even on its own
            if "25" in var and var["25"]["rules"] == 71532:
                assert var["25"][test] != csv["25"], (name, test)


def test_unviable_rules_do_not_depend_on_the_grid_between_19_and_49():
    import csv_provenance as P

    jobs = P.sweep_jobs()
    idx = np.sort(np.random.default_rng(11).choice(len(jobs), 400, replace=False))
    arr = (P.Job * len(idx))(*[P.Job(*jobs[i]) for i in idx])
    Ms = [jobs[i][0] for i in idx]
    cls = {}
    with tempfile.TemporaryDirectory() as tmp:
        lib = P.build_lib(tmp)
        for G in (19, 25):
            res = (C.c_uint32 * (2 * len(idx)))()
            lib.cp_run(0, G, arr, len(idx), 20, 11, 4, res)
            r = np.frombuffer(res, dtype=np.uint32).reshape(-1, 2)
            cls[G] = [[P.classify(int(r[k, col]), Ms[k]) == 0 for k in range(len(idx))] for col in (0, 1)]
    assert cls[19] == cls[25]
    assert 0 < sum(cls[25][0]) < len(idx) and 0 < sum(cls[25][1]) < len(idx)


def test_no_death_threshold_gives_the_csv_25_unviable_count():
    rec = _load("csv_provenance.json")
    csv25 = rec["csv"]["25"]["unviable"]
    for grid in ("25", "49"):
        table = rec["variants"]["v0"][grid]["unviable_if_died_within_M_plus_k"]
        for test in ("current_death", "legacy_death"):
            counts = table[test]
            assert counts["1"] == rec["variants"]["v0"][grid][test]["unviable"]  # the reference's M + 1
            assert csv25 not in counts.values(), (grid, test)
