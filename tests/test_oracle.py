"""Pins of the CPU oracle (oracle/cave_oracle.c) -- no GPU needed.

Each check is chosen so that a plausible slip in the restatement (a dropped
neighbour class, an off-by-one rule bit, a transposed axis, a wrong init box)
fails at least one of them.
"""
import itertools
import math

import numpy as np
import pytest

import oracle as O
import np_model as NP


# ---- rule strings: the cli-rules KAT (src/testing/testRunner.cpp:231-260) ----
def test_cli_rules_kat():
    b = O.rule_string_to_bitmask("3,5-7")
    s = O.rule_string_to_bitmask("2,4")
    assert b == (1 << 2) | (1 << 4) | (1 << 5) | (1 << 6)
    assert s == (1 << 1) | (1 << 3)
    assert O.bitmask_to_rule_string(b) == "3,5-7"
    assert O.bitmask_to_rule_string(s) == "2,4"


def test_rule_string_edge_cases():
    # values outside 1..60 are ignored; whitespace removed; two-element runs stay listed
    assert O.rule_string_to_bitmask(" 0, 61, 1 ") == 1
    assert O.rule_string_to_bitmask("59-70") == (1 << 58) | (1 << 59)
    assert O.bitmask_to_rule_string(0) == "-"
    assert O.bitmask_to_rule_string(0b11) == "1,2"
    assert O.bitmask_to_rule_string(0b111) == "1-3"
    assert O.bitmask_to_rule_string(0b1011) == "1,2,4"


# ---- enumeration (src/common/permutationMath.h:94-128) ----
@pytest.mark.parametrize("N,K,expect", [(6, 2, 21), (14, 2, 105), (26, 2, 351), (26, 4, 17901),
                                        (12, 2, 78), (18, 2, 171)])
def test_valid_bitmask_sizes(N, K, expect):
    v = O.enumerate_valid_bitmasks(N, K)
    assert len(v) == expect == sum(math.comb(N, k) for k in range(1, K + 1))


@pytest.mark.parametrize("N,K", [(6, 2), (12, 3), (8, 8), (5, 0)])
def test_valid_bitmasks_brute_force(N, K):
    brute = [m for m in range(1, 1 << N) if bin(m).count("1") <= K]
    assert O.enumerate_valid_bitmasks(N, K) == brute


def test_sweep_size_matches_csv_total():
    # 2 max-cell-states x subsets {F, E, FE} of "FE" at K = 2 (searchMode.cpp:49-102)
    total = 0
    for M in (2, 3):
        for subset in (1, 2, 3):
            n = O.lib().orc_neighbor_count(subset)
            total += len(O.enumerate_valid_bitmasks(n, 2)) ** 2
    assert total == 71532  # benchmarks/reference/debug/bench_results.csv:2-9 "Total"
    # the K = 0 closed form of permutationMath.h:56-81
    assert O.lib().orc_search_total_permutations(2, 3, 1, 1, 0, 0) == 2 * ((2**6 - 1) ** 2 + (2**12 - 1) ** 2 + (2**18 - 1) ** 2)


# ---- init hash (searchSystem.cpp:244-262) ----
def _wang(s):
    s = ((s ^ 61) ^ (s >> 16)) & 0xFFFFFFFF
    s = (s * 9) & 0xFFFFFFFF
    s ^= s >> 4
    s = (s * 0x27D4EB2D) & 0xFFFFFFFF
    s ^= s >> 15
    return s


def test_wang_hash_vectors():
    for v in (0, 1, 61, 1234, 0xFFFFFFFF, 0x80000000, 73856093):
        assert O.lib().orc_wang_hash(v) == _wang(v)


@pytest.mark.parametrize("p,thr", [(0.35, 1503238553), (0.3, 1288490188), (0.25, 1073741824),
                                   (0.2, 858993459), (1.0, 0xFFFFFFFF), (0.5, 2**31)])
def test_density_threshold(p, thr):
    assert O.lib().orc_density_threshold(p) == thr


@pytest.mark.parametrize("G,S", [(49, 16), (25, 11), (10, 3), (7, 7), (8, 1)])
def test_init_box_and_hash(G, S):
    M, seed, p = 3, 1234, 0.35
    cells = NP.unpack(O.init_cells(G, G, G, M, (S, S, S), 0, p, seed), G, G, G)
    thr = int(p * 2**32)
    lo, hi = (G - S) // 2, (G + S) // 2
    for z, y, x in itertools.product(range(G), repeat=3):
        inside = lo <= x < hi and lo <= y < hi and lo <= z < hi
        h = _wang((x * 73856093 + seed) & 0xFFFFFFFF)
        h = _wang(h ^ ((y * 19349663) & 0xFFFFFFFF))
        h = _wang(h ^ ((z * 83492791) & 0xFFFFFFFF))
        assert cells[z, y, x] == (M if inside and h < thr else 0)
    filled = NP.unpack(O.init_cells(G, G, G, M, (S, S, S), 1, p, seed), G, G, G)
    assert int((filled == M).sum()) == S ** 3


# ---- step ----
FLAG_SETS = [1, 2, 3, 4, 5, 6, 7]


@pytest.mark.parametrize("flags", FLAG_SETS)
@pytest.mark.parametrize("M", [1, 2, 4, 15])
def test_step_matches_independent_model(flags, M):
    rng = np.random.default_rng(flags * 100 + M)
    for (X, Y, Z) in [(1, 1, 1), (2, 2, 2), (3, 1, 1), (9, 3, 2), (13, 5, 7)]:
        cells = rng.integers(0, M + 1, size=(Z, Y, X)).astype(np.uint8)
        birth = int(rng.integers(1, 1 << 26))
        surv = int(rng.integers(1, 1 << 26))
        w0 = (((M - 1) & 0xF) << 60) | birth
        w1 = (flags << 60) | surv
        out, st = O.step(NP.pack(cells), X, Y, Z, w0, w1)
        want, counts = NP.step(cells, M, flags, birth, surv)
        assert np.array_equal(NP.unpack(out, X, Y, Z), want)
        assert st.as_tuple()[:4] == counts
        assert sum(counts) == X * Y * Z


def test_single_cell_birth_counts():
    # a lone live cell with B = {1} births exactly its in-bounds F/E/C neighbours
    for flags, expect in [(1, 6), (2, 12), (4, 8), (3, 18), (7, 26)]:
        X = Y = Z = 5
        cells = np.zeros((Z, Y, X), np.uint8)
        cells[2, 2, 2] = 1
        _, st = O.step(NP.pack(cells), X, Y, Z, (0 << 60) | 1, (flags << 60) | 0)
        assert st.snap.numberOfCellsBorn == expect
        assert st.snap.numberOfDyingCells == 1
    # in a corner only 3 / 7 in-bounds neighbours exist
    cells = np.zeros((3, 3, 3), np.uint8)
    cells[0, 0, 0] = 1
    _, st = O.step(NP.pack(cells), 3, 3, 3, 1, (1 << 60) | 0)
    assert st.snap.numberOfCellsBorn == 3
    _, st = O.step(NP.pack(cells), 3, 3, 3, 1, (7 << 60) | 0)
    assert st.snap.numberOfCellsBorn == 7


def test_dead_grid_stays_dead_and_dying_decrements():
    X, Y, Z = 9, 4, 3
    cells = np.zeros((Z, Y, X), np.uint8)
    out, st = O.step(NP.pack(cells), X, Y, Z, (4 << 60) | ((1 << 26) - 1), (7 << 60) | ((1 << 26) - 1))
    assert st.snap.numberOfDeadCells == X * Y * Z and not NP.unpack(out, X, Y, Z).any()
    cells[:] = 3  # every cell dying (M = 5), nothing live: each decrements by one
    out, st = O.step(NP.pack(cells), X, Y, Z, (4 << 60) | 1, (7 << 60) | 1)
    assert (NP.unpack(out, X, Y, Z) == 2).all() and st.snap.numberOfDyingCells == X * Y * Z


def test_padding_nibbles_stay_zero():
    X, Y, Z = 13, 3, 2
    rng = np.random.default_rng(7)
    cells = rng.integers(0, 2, size=(Z, Y, X)).astype(np.uint8)
    out, _ = O.step(NP.pack(cells), X, Y, Z, 0 | 0x3FFFFFF, (7 << 60) | 0x3FFFFFF)
    w = out.reshape(Z, Y, 2)
    assert ((w[:, :, 1] >> np.uint32(4 * (X - 8))) == 0).all()


def test_count_zero_never_fires():
    # bit 59 set in every mask: count 60 is unreachable, count 0 is never tested
    X = Y = Z = 4
    cells = np.zeros((Z, Y, X), np.uint8)
    out, st = O.step(NP.pack(cells), X, Y, Z, 1 << 59, (7 << 60) | (1 << 59))
    assert st.snap.numberOfCellsBorn == 0


def test_max_rule_bits_index_form():
    # K > 0: low 60 bits are 1-based indices into the valid-bitmask table
    vb = O.enumerate_valid_bitmasks(18, 2)
    rng = np.random.default_rng(3)
    cells = rng.integers(0, 3, size=(6, 7, 11)).astype(np.uint8)
    bi, si = 17, 40
    w0 = (1 << 60) | bi
    w1 = (3 << 60) | si
    a, sa = O.step(NP.pack(cells), 11, 7, 6, w0, w1, vb, 2)
    b, sb = O.step(NP.pack(cells), 11, 7, 6, (1 << 60) | vb[bi - 1], (3 << 60) | vb[si - 1])
    assert np.array_equal(a, b) and sa.as_tuple() == sb.as_tuple()


def test_bbox_and_derived_counts():
    rng = np.random.default_rng(11)
    X, Y, Z, M = 10, 6, 5, 2
    cells = rng.integers(0, M + 1, size=(Z, Y, X)).astype(np.uint8)
    out, st = O.step(NP.pack(cells), X, Y, Z, ((M - 1) << 60) | 0b1100, (3 << 60) | 0b11100)
    o = NP.unpack(out, X, Y, Z)
    zz, yy, xx = np.nonzero(o == M)
    assert st.live == len(xx) == st.snap.numberOfCellsBorn + st.snap.numberOfSurvivingCells
    assert st.changed == st.snap.numberOfCellsBorn + st.snap.numberOfDyingCells
    assert tuple(st.bbox_min) == (xx.min(), yy.min(), zz.min())
    assert tuple(st.bbox_max) == (xx.max(), yy.max(), zz.max())


# ---- manager (searchManager.slang:98-149) and classification ----
def test_manager_state_machine():
    L = O.lib()
    info, perm = O.GridInfo(), O.Permutation()
    perm.birthAndMaxCellStateRules = (2 << 60) | 1
    perm.survivalAndNeighborhoodRules = (3 << 60) | 1
    ci, cp = O.GridInfo(), O.Permutation()
    snap = O.Snapshot(10, 2, 3, 4)
    st = L.orc_manager_tick(info, snap, perm, 5, 5, 3, ci, cp)
    assert st == 0 and ci.numberOfSimulationTicksSurvived == 1 and ci.totalNumberOfCellsThatDied == 10
    assert ci.totalNumberOfCellsThatWereBorn == 3 and ci.finalAliveCount == 4 and ci.maxAliveCount == 4
    L.orc_manager_tick(info, O.Snapshot(1, 1, 0, 9), perm, 5, 5, 3, ci, cp)
    st = L.orc_manager_tick(info, O.Snapshot(1, 1, 1, 2), perm, 5, 5, 3, ci, cp)
    assert st == 2 and ci.maxAliveCount == 9 and ci.finalAliveCount == 2
    # the permutation advanced (birth first) and GridInfo was reset
    assert perm.birthAndMaxCellStateRules == (2 << 60) | 2 and info.numberOfSimulationTicksSurvived == 0
    # early death: no survivor and no birth in one tick
    st = L.orc_manager_tick(info, O.Snapshot(5, 5, 0, 0), perm, 5, 5, 3, ci, cp)
    assert st == 1
    # classification
    g = O.GridInfo()
    g.status, g.numberOfSimulationTicksSurvived = 2, 1
    assert L.orc_classify(g, 3) == 2
    g.status, g.numberOfSimulationTicksSurvived = 1, 5
    assert L.orc_classify(g, 3) == 1
    g.numberOfSimulationTicksSurvived = 4
    assert L.orc_classify(g, 3) == 0


def test_manager_exhaustion_order():
    # birth index runs fastest; at its end survival advances and birth resets to 1
    L = O.lib()
    info, perm, ci, cp = O.GridInfo(), O.Permutation(), O.GridInfo(), O.Permutation()
    perm.birthAndMaxCellStateRules = (1 << 60) | 1
    perm.survivalAndNeighborhoodRules = (2 << 60) | 1
    seen = []
    for _ in range(10):
        seen.append((perm.birthAndMaxCellStateRules & 0xF, perm.survivalAndNeighborhoodRules & 0xF))
        L.orc_manager_tick(info, O.Snapshot(0, 0, 0, 0), perm, 2, 2, 20, ci, cp)
        if perm.birthAndMaxCellStateRules == 0:
            break
    assert seen == [(1, 1), (2, 1), (1, 2), (2, 2)]
    assert perm.survivalAndNeighborhoodRules == 0


# ---- a whole chunk: the reference's published integration counts ----
@pytest.mark.slow
def test_search_counts_25_match_fixture():
    """The oracle's full 25^3 sweep (71,532 rule evaluations) reproduces the
    committed counts; those counts in turn pin the oracle to the reference's
    CSV where the CSV's semantics are known (see test_csv_pins_*)."""
    import json, os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "search_counts.json")))
    n, (unv, pos, via) = O.run_search(25, 3, 2, 2, 3, 20, 11, 1, 0.5, 0, 8)
    want = fx["oracle"]["25"]
    assert (n, unv, pos, via) == (want["total"], want["unviable"], want["possible"], want["viable"])


def test_csv_pins_unviable_count_under_legacy_death():
    """benchmarks/reference/debug/bench_results.csv:6-9 (49^3) was produced
    before the early-death fix of searchManager.slang:121-127: under the
    superseded death test (shaders/source/search/manager.comp:115) the oracle
    reproduces the CSV's unviable count exactly and its viable + possible
    total exactly; 248 of the 46,800 long-lived rules (0.5%) split differently
    between viable and possible (parity unpinned for that split)."""
    import json, os
    fx = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "search_counts.json")))
    csv, orc = fx["csv"]["49"], fx["oracle"]["49_legacy_death"]
    assert orc["total"] == csv["total"] == 71532
    assert orc["unviable"] == csv["unviable"]
    assert orc["viable"] + orc["possible"] == csv["viable"] + csv["possible"]
    # the early-dying rules do not reach the walls, so 25^3 agrees with 49^3 there
    assert fx["oracle"]["25_legacy_death"]["unviable"] == csv["unviable"]


def test_run_rule_evaluates_all_zero_first_word():
    # M = 1 with no birth count packs to an all-zero first rule word, the
    # reference's exhausted-slot marker; as an explicit rule it still runs:
    # nothing is born, so with no survival count everything dies in tick 1
    init = O.init_cells(9, 9, 9, 1, (5, 5, 5), 0, 0.5, 3)
    t, info, rows, _ = O.run_rule(9, 9, 9, 1, 1, 0, 0, 40, init)
    assert t == 1 and info.status == 1 and rows[0].snap.numberOfCellsBorn == 0
    # survival of every count keeps the live cells with a live neighbour
    t, info, _, _ = O.run_rule(9, 9, 9, 1, 1, 0, 0x3F, 40, init)
    assert info.numberOfSimulationTicksSurvived == t and t >= 1


# ---- the headless render-path step (rayMarchCompute.slang) and init (cellInit.slang) ----
@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 1), (3, 2, 2), (5, 4, 3), (9, 7, 6), (17, 5, 4)])
@pytest.mark.parametrize("flags", [1, 2, 3, 4, 5, 7])
def test_render_step_wraps_like_a_torus(dims, flags):
    X, Y, Z = dims
    rng = np.random.default_rng(X * 100 + Y * 10 + Z + flags)
    for M in (1, 3, 15):
        cells = rng.integers(0, M + 1, size=(Z, Y, X)).astype(np.uint8)
        birth, surv = int(rng.integers(1, 1 << 26)), int(rng.integers(1, 1 << 26))
        w0 = ((M - 1) & 0xF) << 60 | birth
        w1 = (flags | 8) << 60 | surv
        out, st = O.step_render(NP.pack(cells), X, Y, Z, w0, w1)
        want, counts = NP.step(cells, M, flags, birth, surv, wrap=True)
        assert np.array_equal(NP.unpack(out, X, Y, Z), want)
        assert st.as_tuple()[:4] == counts
        # without W the render step is the search step
        w1n = flags << 60 | surv
        a, _ = O.step_render(NP.pack(cells), X, Y, Z, w0, w1n)
        b, _ = O.step(NP.pack(cells), X, Y, Z, w0, w1n)
        assert np.array_equal(a, b)


def test_render_step_self_counts_on_one_cell_torus():
    # a 1^3 torus: all 26 offsets reach the cell itself
    cells = np.array([[[1]]], dtype=np.uint8)
    for flags, n in ((1, 6), (2, 12), (4, 8), (7, 26)):
        w0 = 0 << 60 | (1 << (n - 1))  # M = 1, born on n
        w1 = (flags | 8) << 60 | (1 << (n - 1))  # survives on exactly n
        out, st = O.step_render(NP.pack(cells), 1, 1, 1, w0, w1)
        assert st.snap.numberOfSurvivingCells == 1


@pytest.mark.parametrize("p,thr", [(0.35, 1503238528), (0.5, 2147483648), (0.25, 1073741824),
                                   (0.2, 858993472), (0.3, 1288490240), (0.0, 0)])
def test_render_density_threshold(p, thr):
    # uint(float(p) * 2^32) after the "%f" round trip; the search path's double
    # floor(p * 2^32) differs for p = 0.35 (1503238553), 0.2 (858993459), 0.3
    t, all_alive = O.render_density_threshold(p)
    assert (t, all_alive) == (thr, False)
    f = np.float32(float("%f" % np.float32(p)))
    assert int(np.float32(f) * np.float32(2.0**32)) == thr


def test_render_density_threshold_round_trip_and_full():
    # 1/3 becomes "0.333333" first
    t, _ = O.render_density_threshold(1.0 / 3.0)
    assert t == int(np.float32("0.333333") * np.float32(2.0**32))
    assert O.render_density_threshold(1.0) == (0, True)
    assert O.render_density_threshold(0.9999999)[1]  # "%f" rounds it to 1.000000


def test_render_init_is_unsalted_hash_with_float_threshold():
    X, Y, Z, M, spawn = 23, 19, 17, 5, (12, 9, 11)
    for p in (0.35, 0.5):
        got = NP.unpack(O.init_cells_render(X, Y, Z, M, spawn, 0, p), X, Y, Z)
        thr, _ = O.render_density_threshold(p)
        want = np.zeros((Z, Y, X), np.uint8)
        for z in range((Z - spawn[2]) // 2, (Z + spawn[2]) // 2):
            for y in range((Y - spawn[1]) // 2, (Y + spawn[1]) // 2):
                for x in range((X - spawn[0]) // 2, (X + spawn[0]) // 2):
                    if O.lib().orc_cell_hash(x, y, z, 0) < thr:
                        want[z, y, x] = M
        assert np.array_equal(got, want)
    # where the two thresholds agree, the render init is the search init with seed 0
    a = O.init_cells_render(X, Y, Z, M, spawn, 0, 0.5)
    b = O.init_cells(X, Y, Z, M, spawn, 0, 0.5, 0)
    assert np.array_equal(a, b)


# ---- ERD shape (searchSimulationERD.slang), pinned by the honeycomb's geometry ----
@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 3, 2), (5, 4, 3), (9, 7, 6), (13, 8, 5)])
@pytest.mark.parametrize("flags", [1, 2, 3, 9, 10, 11])
def test_erd_step_matches_geometric_model(dims, flags):
    X, Y, Z = dims
    rng = np.random.default_rng(X * 1000 + Y * 10 + Z + flags)
    N = O.lib().orc_erd_neighbor_count(flags)
    for M in (1, 4, 15):
        cells = rng.integers(0, M + 1, size=(Z, Y, X)).astype(np.uint8)
        birth, surv = int(rng.integers(1, 1 << N)), int(rng.integers(1, 1 << N))
        w0 = ((M - 1) & 0xF) << 60 | birth
        w1 = flags << 60 | surv
        out, st = O.step_erd(NP.pack_linear(cells), X, Y, Z, w0, w1)
        want, counts = NP.erd_step(cells, M, flags & 3, birth, surv, wrap=bool(flags & 8))
        assert np.array_equal(NP.unpack_linear(out, X, Y, Z), want)
        assert st.as_tuple()[:4] == counts


def test_erd_neighbour_relation_is_symmetric_and_complete():
    # a lone live cell with B = {1} births exactly its 12 / 4 / 16 neighbours,
    # and every cell it births births it back (faces are shared)
    X = Y = Z = 9
    for flags, n in ((1, 12), (2, 4), (3, 16)):
        for cy in (4, 5):
            cells = np.zeros((Z, Y, X), np.uint8)
            cells[4, cy, 4] = 1
            out, st = O.step_erd(NP.pack_linear(cells), X, Y, Z, 0 << 60 | 1, flags << 60 | 0)
            born = np.argwhere(NP.unpack_linear(out, X, Y, Z) == 1)
            assert st.snap.numberOfCellsBorn == n == len(born)
            for z, y, x in born:
                back = np.zeros((Z, Y, X), np.uint8)
                back[z, y, x] = 1
                o2, _ = O.step_erd(NP.pack_linear(back), X, Y, Z, 0 << 60 | 1, flags << 60 | 0)
                assert NP.unpack_linear(o2, X, Y, Z)[4, cy, 4] == 1


def test_erd_init_and_hash():
    X, Y, Z, M = 11, 9, 7, 3
    a = O.init_cells_erd(X, Y, Z, M, (6, 5, 4), 0, 0.4, 17)
    b = O.init_cells(X, Y, Z, M, (6, 5, 4), 0, 0.4, 17)
    # same cells, linear vs row-padded layout; the same canonical hash
    assert np.array_equal(NP.unpack_linear(a, X, Y, Z), NP.unpack(b, X, Y, Z))
    assert O.erd_hash(a, X, Y, Z) == O.canonical_hash(b, X, Y, Z)


def test_erd_search_counts_cover_the_rule_space():
    n, (u, p, v) = O.run_search_shape(7, 1, 3, 2, 2, 2, 6, 3, 1, 0.5, 0, 8)
    # F (12), E (4), FE (16) slots at K = 2
    assert n == sum((N + N * (N - 1) // 2) ** 2 for N in (12, 4, 16)) == u + p + v
