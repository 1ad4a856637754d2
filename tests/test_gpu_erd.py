"""The ERD cell shape (searchSimulationERD.slang) on the GPU vs the CPU oracle:
the step state by state (with and without W), batches mixing ERD and cube
simulations, the search sweep's counts and the CLI's --shape ed."""
import json
import os
import subprocess

import numpy as np
import pytest

import cave_amd as cv
import np_model as NP
import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
CLI = os.path.join(os.path.dirname(HERE), "cave_amd", "_lib", "cave")

CASES = [  # (X, Y, Z, M, flags)
    (1, 1, 1, 1, 3), (2, 3, 2, 2, 11), (5, 4, 3, 4, 1), (9, 7, 6, 1, 9), (13, 8, 5, 15, 3),
    (33, 12, 9, 3, 2), (40, 17, 11, 2, 11), (64, 5, 31, 5, 10),
]


@pytest.mark.parametrize("X,Y,Z,M,flags", CASES)
def test_erd_step_state_by_state(X, Y, Z, M, flags):
    rng = np.random.default_rng(X * 77 + Y * 5 + Z + M + flags)
    N = O.lib().orc_erd_neighbor_count(flags)
    birth, surv = int(rng.integers(1, 1 << N)), int(rng.integers(1, 1 << N))
    cells = NP.pack_linear(rng.integers(0, M + 1, size=(Z, Y, X)).astype(np.uint8))
    w0 = ((M - 1) & 0xF) << 60 | birth
    w1 = flags << 60 | surv
    print(f"erd case {X}x{Y}x{Z} M={M} flags={flags}", flush=True)
    with cv.Step((X, Y, Z), M, flags, birth, surv, initial_cells=cells, shape=cv.SHAPE_ERD) as h:
        assert h.kernel == cv.KERNEL_ERD
        ref = cells
        for t in range(6):
            ref, st = O.step_erd(ref, X, Y, Z, w0, w1)
            (got,), _ = h.run(1)
            assert got.as_tuple() == st.as_tuple(), f"tick {t}"
            assert np.array_equal(h.cells(), ref), f"tick {t}"


def test_erd_init_on_device():
    X, Y, Z, M = 45, 20, 13, 6
    with cv.Step((X, Y, Z), M, 3, 1, 1, spawn=(30, 10, 9), density=0.4, seed=5, shape=cv.SHAPE_ERD) as h:
        assert np.array_equal(h.cells(), O.init_cells_erd(X, Y, Z, M, (30, 10, 9), 0, 0.4, 5))


def test_mixed_erd_and_cube_batch():
    rng = np.random.default_rng(3)
    sims = []
    for i in range(16):
        erd = i % 2 == 0
        dims = (int(rng.integers(8, 40)), int(rng.integers(4, 30)), int(rng.integers(4, 20)))
        flags = int((1, 2, 3, 11)[i % 4]) if erd else int((1, 3, 7, 15)[i % 4])
        N = O.lib().orc_erd_neighbor_count(flags) if erd else O.lib().orc_neighbor_count(flags & 7)
        sims.append(cv.Sim(dims=dims, M=int(rng.integers(1, 8)), flags=flags, ticks=int(rng.integers(4, 16)),
                           birth=int(rng.integers(1, 1 << N)), survival=int(rng.integers(1, 1 << N)),
                           spawn=tuple(max(1, d // 2) for d in dims), spawn_mode=cv.SPAWN_RANDOM, density=0.35,
                           seed=int(rng.integers(0, 1 << 20)),
                           sim_flags=cv.SIM_FINAL_HASH | (cv.SIM_SHAPE_ERD if erd else 0)))
    res, rows, _ = cv.batch_run(sims, per_tick=True)
    for i, (s, r) in enumerate(zip(sims, res)):
        X, Y, Z = s.dims
        if s.sim_flags & cv.SIM_SHAPE_ERD:
            init = O.init_cells_erd(X, Y, Z, s.M, s.spawn, s.spawn_mode, s.density, s.seed)
            t, info, orows, final = O.run_rule_erd(X, Y, Z, s.M, s.flags, s.birth, s.survival, s.ticks, init)
            fh = O.erd_hash(final, X, Y, Z)
        else:
            # the cube search kernel ignores W (searchSimulationCube.slang:26-29)
            init = O.init_cells(X, Y, Z, s.M, s.spawn, s.spawn_mode, s.density, s.seed)
            t, info, orows, final = O.run_rule(X, Y, Z, s.M, s.flags & 7, s.birth, s.survival, s.ticks, init)
            fh = O.canonical_hash(final, X, Y, Z)
        assert r.ticks_run == t and r.info.finalAliveCount == info.finalAliveCount, f"sim {i}"
        assert [x.as_tuple() for x in rows[i]] == [x.as_tuple() for x in orows], f"sim {i}"
        assert r.final_hash == fh, f"sim {i}"


@pytest.mark.parametrize("wrap", [False, True], ids=["plain", "wrap"])
def test_erd_search_counts(wrap):
    flags = cv.FACES | cv.EDGES | (cv.WRAP if wrap else 0)
    s = cv.Search(7, flags, 2, 2, 2, 6, spawn=3, spawn_mode=cv.SPAWN_FILLED, shape=cv.SHAPE_ERD)
    s.run()
    res = s.drain()
    p = s.progress()
    per = sum((N + N * (N - 1) // 2) ** 2 for N in (12, 4, 16))
    assert p["total"] == len(res) == per * (2 if wrap else 1)
    # the plain slots agree with the oracle's ERD sweep
    n, (u, q, v) = O.run_search_shape(7, 1, 3, 2, 2, 2, 6, 3, 1, 0.5, 0, 8)
    plain = [r for r in res if not (r.permutation.survivalAndNeighborhoodRules >> 63)]
    assert len(plain) == n
    assert (sum(r.category == 0 for r in plain), sum(r.category == 1 for r in plain),
            sum(r.category == 2 for r in plain)) == (u, q, v)
    assert all(r.permutation.shapeAndGridConfiguration == 0 for r in res)
    if wrap:
        init = O.init_cells_erd(7, 7, 7, 2, (3, 3, 3), 1, 0.5, 0)
        wr = [r for r in res if r.permutation.survivalAndNeighborhoodRules >> 63]
        for r in wr[:: max(1, len(wr) // 40)]:
            fl = (r.permutation.survivalAndNeighborhoodRules >> 60) & 0xF
            b = r.permutation.birthAndMaxCellStateRules & ((1 << 60) - 1)
            sv = r.permutation.survivalAndNeighborhoodRules & ((1 << 60) - 1)
            t, info, _, _ = O.run_rule_erd(7, 7, 7, 2, fl, b, sv, 6, init)
            assert (r.ticksSurvived, r.finalAliveCount, r.totalBirths) == (
                info.numberOfSimulationTicksSurvived, info.finalAliveCount, info.totalNumberOfCellsThatWereBorn)


def test_cli_shape_ed(tmp_path):
    out = subprocess.run([CLI, "--mode", "search", "--shape", "ED", "--grid", "7", "--neighborhood", "FE",
                          "--chunks-per-config", "2", "--max-rule-bits", "1", "--min-cs", "2", "--max-cs-range", "2",
                          "--ticks", "6", "--output", str(tmp_path)], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    folder = tmp_path / "ElongatedRhombicDodecahedron"
    files = sorted(p.name for p in folder.iterdir())
    assert any(f.endswith("_viable.jsonl") for f in files)
    lines = [json.loads(l) for f in files if f.endswith(".jsonl") and "aggregate" not in f for l in open(folder / f)]
    assert len(lines) == sum(N * N for N in (12, 4, 16))  # K = 1: N masks per side


def erd_sims(cases, seed, sim_flags=cv.SIM_FINAL_HASH):
    rng = np.random.default_rng(seed)
    sims = []
    for dims, M, flags, ticks in cases:
        N = O.lib().orc_erd_neighbor_count(flags)
        birth = sum(1 << b for b in range(N) if rng.random() < 0.25) or 1
        surv = sum(1 << b for b in range(N) if rng.random() < 0.3) or 1
        sims.append(cv.Sim(dims=dims, M=M, flags=flags, ticks=ticks, birth=birth, survival=surv,
                           spawn=tuple(max(1, d // 2) for d in dims), spawn_mode=cv.SPAWN_RANDOM, density=0.3,
                           seed=int(rng.integers(0, 1 << 20)), sim_flags=sim_flags | cv.SIM_SHAPE_ERD))
    return sims


# (dims, M, flags, ticks): LDS-resident grids, then grids in global memory
# (above the LDS share of a 1024-thread block), odd and even row counts
ERD_BATCH = [((9, 7, 6), 1, 1, 12), ((24, 13, 10), 3, 2, 15), ((40, 21, 17), 5, 3, 20), ((33, 32, 31), 2, 3, 18),
             ((100, 81, 60), 4, 3, 16), ((70, 64, 50), 7, 1, 14), ((96, 45, 66), 2, 2, 12)]


@pytest.mark.parametrize("max_parts", [1, 2, 5], ids=["whole", "2parts", "5parts"])
def test_erd_persistent_engine_matches_oracle(max_parts):
    """ERD grids on the persistent batch kernel (its bit-sliced ERD tick), in
    LDS and in global memory, whole and split over blocks: every per-tick
    row, the grid info and the final hash against the oracle's
    orc_run_rule_erd (searchSimulationERD.slang:59-217)."""
    sims = erd_sims(ERD_BATCH, 11)
    print(f"erd persistent max_parts={max_parts}", flush=True)
    res, rows, _ = cv.batch_run(sims, engine=cv.ENGINE_PERSISTENT, per_tick=True, max_parts=max_parts)
    for i, (s, r) in enumerate(zip(sims, res)):
        X, Y, Z = s.dims
        init = O.init_cells_erd(X, Y, Z, s.M, s.spawn, s.spawn_mode, s.density, s.seed)
        t, info, orows, final = O.run_rule_erd(X, Y, Z, s.M, s.flags, s.birth, s.survival, s.ticks, init)
        assert r.ticks_run == t, f"sim {i}"
        for f, _ in O.GridInfo._fields_:
            assert getattr(r.info, f) == getattr(info, f), (i, f)
        assert [x.as_tuple() for x in rows[i]] == [x.as_tuple() for x in orows], f"sim {i}"
        assert r.final_hash == O.erd_hash(final, X, Y, Z), f"sim {i}"


def test_erd_persistent_equals_per_config_engine():
    """Larger ERD grids (up to 150^3, 60 ticks, no early exit) on the
    persistent kernel and on the per-config engine (the reference-form ERD
    kernel): same results and final cells."""
    rng = np.random.default_rng(5)
    cases = [((int(e), int(e), int(e)), int(rng.integers(1, 16)), int((1, 2, 3)[i % 3]), 60)
             for i, e in enumerate((48, 64, 90, 121, 150, 77))]
    sims = erd_sims(cases, 6, sim_flags=cv.SIM_FINAL_HASH | cv.SIM_NO_EARLY_EXIT)
    a, _, _ = cv.batch_run(sims, engine=cv.ENGINE_PERSISTENT)
    b, _, _ = cv.batch_run(sims, engine=cv.ENGINE_PLAIN)
    for i in range(len(sims)):
        assert a[i].ticks_run == b[i].ticks_run == 60
        assert bytes(a[i].info) == bytes(b[i].info), f"sim {i}"
        assert a[i].last.as_tuple() == b[i].last.as_tuple(), f"sim {i}"
        assert a[i].final_hash == b[i].final_hash, f"sim {i}"


def test_server_refuses_erd_requests():
    s = erd_sims([((8, 8, 8), 2, 3, 4)], 1)[0]
    with cv.Server() as srv:
        with pytest.raises(cv.CaveError) as e:
            srv.submit(s, 1)
    assert e.value.code == cv.CAVE_E_UNSUPPORTED


# wrapping ERD grids (W = 8): tiny axes (the same cell counts once per offset
# that reaches it), rows of one word, rows ending on a word boundary, the LDS
# and the global tier
ERD_WRAP_BATCH = [((1, 1, 1), 1, 11, 6), ((2, 3, 2), 2, 9, 8), ((3, 2, 5), 3, 10, 8), ((9, 7, 6), 4, 11, 10),
                  ((20, 5, 7), 2, 9, 10), ((64, 20, 12), 5, 11, 12), ((33, 32, 31), 3, 10, 12),
                  ((40, 21, 17), 7, 11, 14), ((100, 81, 60), 2, 11, 10), ((96, 64, 40), 6, 9, 9)]


@pytest.mark.parametrize("max_parts", [1, 3], ids=["whole", "3parts"])
def test_erd_wrap_persistent_engine_matches_oracle(max_parts):
    """Wrapping ERD grids on the persistent kernel's wrap form: rows y+-1 and
    planes z+-1 past a face come from the opposite face, and x-neighbours
    across the row ends from the row's other end; every per-tick row, grid
    info and final hash against the oracle (which wraps every coordinate of
    every offset, searchSimulationERD.slang with W)."""
    sims = erd_sims(ERD_WRAP_BATCH, 12)
    print(f"erd wrap persistent max_parts={max_parts}", flush=True)
    res, rows, _ = cv.batch_run(sims, engine=cv.ENGINE_PERSISTENT, per_tick=True, max_parts=max_parts)
    for i, (s, r) in enumerate(zip(sims, res)):
        X, Y, Z = s.dims
        init = O.init_cells_erd(X, Y, Z, s.M, s.spawn, s.spawn_mode, s.density, s.seed)
        t, info, orows, final = O.run_rule_erd(X, Y, Z, s.M, s.flags, s.birth, s.survival, s.ticks, init)
        assert r.ticks_run == t, f"sim {i}"
        for f, _ in O.GridInfo._fields_:
            assert getattr(r.info, f) == getattr(info, f), (i, f)
        assert [x.as_tuple() for x in rows[i]] == [x.as_tuple() for x in orows], f"sim {i}"
        assert r.final_hash == O.erd_hash(final, X, Y, Z), f"sim {i}"
