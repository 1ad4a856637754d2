"""The rule aggregate (cave_aggregate_*, RuleAggregator of
src/common/ruleAggregator.h:24-135) on hand-built results, no device work:
grouping by rule, the run counts per category, min / max / average ticks, the
viability rate, the line format and the batch form's seed folding."""
import json

import numpy as np

import cave_amd as cv


def test_record_and_line_format(tmp_path):
    a = cv.Aggregate()
    p = cv.encode_params(3, cv.FACES | cv.EDGES, 0b110000, 0b111000)
    for ticks, cat in ((5, 2), (7, 0), (9, 1)):
        a.record(p, ticks, cat)
    q = cv.encode_params(1, cv.FACES, 0b1, 0b10)
    a.record(q, 4, 0)
    assert len(a) == 2
    path = str(tmp_path / "agg.jsonl")
    a.write_jsonl(path)
    lines = open(path).read().splitlines()
    # keys in the reference's order; floats as the default ostream writes them
    # (six significant digits: 1/3 -> 0.333333, 21/3 -> 7)
    assert lines == [
        '{"birth":1,"survival":2,"maxCellState":1,"neighborhoodFlags":1,"shapeId":0,"totalRuns":1,"viableRuns":0,'
        '"possibleRuns":0,"unviableRuns":1,"viabilityRate":0,"minTicksSurvived":4,"maxTicksSurvived":4,'
        '"avgTicksSurvived":4}',
        '{"birth":48,"survival":56,"maxCellState":3,"neighborhoodFlags":3,"shapeId":0,"totalRuns":3,"viableRuns":1,'
        '"possibleRuns":1,"unviableRuns":1,"viabilityRate":0.333333,"minTicksSurvived":5,"maxTicksSurvived":9,'
        '"avgTicksSurvived":7}',
    ]
    a.close()


def test_batch_folds_seeds_into_one_record_per_rule(tmp_path):
    # 3 rule configs x 4 seeds, and a fourth config repeating the first rule:
    # RuleAggregator keys on the rule, so that one folds into the first record
    rules = [(2, 3, 0b1100, 0b11000), (4, 7, 0b11, 0b101), (1, 1, 0b10, 0b10), (2, 3, 0b1100, 0b11000)]
    seeds = 4
    specs = np.zeros(len(rules) * seeds, dtype=cv.SIM_DTYPE)
    res = np.zeros(len(specs), dtype=cv.RESULT_DTYPE)
    for c, (M, fl, b, s) in enumerate(rules):
        for k in range(seeds):
            i = c * seeds + k
            specs[i]["M"], specs[i]["flags"], specs[i]["birth"], specs[i]["survival"] = M, fl, b, s
            res[i]["ticks_survived"] = 10 + i
            res[i]["classification"] = i % 3
    a = cv.Aggregate()
    a.record_batch(specs, res)
    assert len(a) == 3
    path = str(tmp_path / "b.jsonl")
    a.write_jsonl(path)
    recs = {(r["birth"], r["survival"], r["maxCellState"]): r for r in map(json.loads, open(path))}
    r = recs[(0b1100, 0b11000, 2)]
    rows = list(range(0, 4)) + list(range(12, 16))  # configs 0 and 3
    assert r["totalRuns"] == 8 and r["neighborhoodFlags"] == 3
    assert r["viableRuns"] == sum(1 for i in rows if i % 3 == 2)
    assert r["possibleRuns"] == sum(1 for i in rows if i % 3 == 1)
    assert r["unviableRuns"] == sum(1 for i in rows if i % 3 == 0)
    assert r["minTicksSurvived"] == 10 and r["maxTicksSurvived"] == 25
    assert abs(r["avgTicksSurvived"] - sum(10 + i for i in rows) / 8) < 1e-4
    assert recs[(0b11, 0b101, 4)]["totalRuns"] == 4
    a.close()


def test_empty_aggregate_writes_an_empty_file(tmp_path):
    a = cv.Aggregate()
    path = str(tmp_path / "e.jsonl")
    a.write_jsonl(path)
    assert open(path).read() == "" and len(a) == 0
