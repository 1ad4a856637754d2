"""Batched evaluation (cave_batch_run) vs the CPU oracle's per-rule runs.

Every simulation is compared field by field: the manager's GridInfo, the
classification, the per-tick counters / bbox and the FNV-1a hash of the final
cells.  Both engines (persistent and plain per-config) are checked, and both
storage tiers of the persistent kernel (LDS-resident and global arena)."""
import numpy as np
import pytest

import cave_amd as cv
import oracle as O

pytestmark = pytest.mark.gpu


def oracle_run(sim: cv.Sim):
    X, Y, Z = sim.dims
    init = O.init_cells(X, Y, Z, sim.M, sim.spawn, sim.spawn_mode, sim.density, sim.seed)
    if sim.sim_flags & cv.SIM_NO_EARLY_EXIT:
        # step for exactly `ticks` ticks, manager semantics per tick
        cells, rows = init, []
        w0 = ((sim.M - 1) & 0xF) << 60 | sim.birth
        w1 = sim.flags << 60 | sim.survival
        for _ in range(sim.ticks):
            cells, st = O.step(cells, X, Y, Z, w0, w1)
            rows.append(st)
        return None, rows, cells
    t, info, rows, final = O.run_rule(X, Y, Z, sim.M, sim.flags, sim.birth, sim.survival, sim.ticks, init)
    return info, rows, final


def random_sims(n, seed, dims_choices, ticks=12, flags_choices=(1, 2, 3, 4, 5, 6, 7), p=0.2):
    rng = np.random.default_rng(seed)
    sims = []
    for i in range(n):
        dims = tuple(int(v) for v in dims_choices[i % len(dims_choices)])
        flags = int(flags_choices[i % len(flags_choices)])
        M = int(rng.integers(1, 16))
        N = O.lib().orc_neighbor_count(flags)
        birth = sum(1 << b for b in range(N) if rng.random() < p) or 1
        surv = sum(1 << b for b in range(N) if rng.random() < p) or 2
        spawn = tuple(max(1, d // 2) for d in dims)
        sims.append(cv.Sim(dims=dims, M=M, flags=flags, ticks=ticks, birth=birth, survival=surv,
                           spawn=spawn, spawn_mode=cv.SPAWN_RANDOM, density=0.3,
                           seed=int(rng.integers(0, 2**31)), sim_flags=cv.SIM_FINAL_HASH))
    return sims


def check(sims, engine, per_tick=True):
    print(f"batch of {len(sims)} engine={engine}", flush=True)
    res, rows, _ = cv.batch_run(sims, engine=engine, per_tick=per_tick)
    for i, (s, r) in enumerate(zip(sims, res)):
        info, orows, final = oracle_run(s)
        X, Y, Z = s.dims
        if info is not None:
            got = (r.info.totalNumberOfCellsThatDied, r.info.totalNumberOfCellsThatWereBorn,
                   r.info.numberOfSimulationTicksSurvived, r.info.status, r.info.maxAliveCount,
                   r.info.finalAliveCount)
            want = (info.totalNumberOfCellsThatDied, info.totalNumberOfCellsThatWereBorn,
                    info.numberOfSimulationTicksSurvived, info.status, info.maxAliveCount,
                    info.finalAliveCount)
            assert got == want, f"sim {i} {s}"
            assert r.classification == O.lib().orc_classify(info, s.M), f"sim {i}"
        assert r.ticks_run == len(orows), f"sim {i}"
        if per_tick:
            assert [t.as_tuple() for t in rows[i]] == [t.as_tuple() for t in orows], f"sim {i}"
        assert r.last.as_tuple() == orows[-1].as_tuple(), f"sim {i}"
        assert r.final_hash == O.canonical_hash(final, X, Y, Z), f"sim {i}"
    return res


@pytest.mark.parametrize("engine", [cv.ENGINE_PERSISTENT, cv.ENGINE_PLAIN])
def test_small_grids_lds_tier(engine):
    dims = [(1, 1, 1), (7, 5, 3), (33, 9, 4), (49, 12, 10), (16, 16, 16), (40, 3, 30), (64, 20, 2)]
    check(random_sims(21, 1, dims), engine)


@pytest.mark.parametrize("engine", [cv.ENGINE_PERSISTENT, cv.ENGINE_PLAIN])
def test_global_tier_grids(engine):
    # storage above the LDS budget: 6 planes x ceil(X/32) x Y x Z words > 159 KB
    dims = [(70, 40, 30), (100, 33, 21), (65, 64, 64)]
    check(random_sims(6, 2, dims, ticks=6), engine)


def test_engines_agree_on_uniform_search_slice():
    # 49^3 FE, C in 2..5, <= 2 birth / survival bits from counts 1..18, 12^3 Bernoulli(0.3)
    vb = cv.enumerate_valid_bitmasks(18, 2)
    rng = cv.Rng(11)
    sims = []
    for c in range(6):
        M = 1 + rng.next() % 4
        b, s = vb[rng.next() % len(vb)], vb[rng.next() % len(vb)]
        for seed in range(3):
            sims.append(cv.Sim(dims=(49, 49, 49), M=M, flags=3, ticks=20, birth=b, survival=s,
                               spawn=(12, 12, 12), spawn_mode=cv.SPAWN_RANDOM, density=0.3, seed=seed,
                               sim_flags=cv.SIM_FINAL_HASH))
    a = check(sims, cv.ENGINE_PERSISTENT)
    b, _, _ = cv.batch_run(sims, engine=cv.ENGINE_PLAIN)
    assert [x.final_hash for x in a] == [x.final_hash for x in b]


def test_no_early_exit_runs_every_tick():
    sims = random_sims(6, 3, [(20, 20, 20), (33, 17, 9)], ticks=15)
    for s in sims:
        s.sim_flags |= cv.SIM_NO_EARLY_EXIT
    res, rows, _ = cv.batch_run(sims, per_tick=True)
    for s, r, rr in zip(sims, res, rows):
        _, orows, final = oracle_run(s)
        assert r.ticks_run == s.ticks
        assert [t.as_tuple() for t in rr] == [t.as_tuple() for t in orows]
        assert r.final_hash == O.canonical_hash(final, *s.dims)


def test_many_jobs_work_queue():
    # more jobs than resident blocks: every result slot is filled exactly once
    sims = random_sims(3000, 4, [(24, 24, 24), (12, 30, 17), (40, 40, 8)], ticks=5)
    res, _, _ = cv.batch_run(sims, per_tick=False)
    pick = list(range(0, 3000, 97))
    for i in pick:
        info, orows, final = oracle_run(sims[i])
        assert res[i].info.numberOfSimulationTicksSurvived == info.numberOfSimulationTicksSurvived
        assert res[i].final_hash == O.canonical_hash(final, *sims[i].dims)


def test_split_simulations_match_whole():
    # global-tier grids split over 2..16 blocks (z-slabs, a rendezvous per tick)
    # give the same results, per-tick rows and final cells as one block per grid
    dims = [(70, 40, 30), (100, 33, 21), (65, 64, 64), (96, 50, 37)]
    sims = random_sims(8, 5, dims, ticks=10)
    for s in sims[::2]:
        s.sim_flags |= cv.SIM_NO_EARLY_EXIT
    print("whole", flush=True)
    whole, wrows, _ = cv.batch_run(sims, per_tick=True, max_parts=1)
    for parts in (2, 3, 7, 16):
        print(f"max_parts={parts}", flush=True)
        split, srows, _ = cv.batch_run(sims, per_tick=True, max_parts=parts)
        for i in range(len(sims)):
            assert split[i].ticks_run == whole[i].ticks_run, f"sim {i} parts {parts}"
            assert split[i].last.as_tuple() == whole[i].last.as_tuple(), f"sim {i} parts {parts}"
            assert split[i].final_hash == whole[i].final_hash, f"sim {i} parts {parts}"
            assert [t.as_tuple() for t in srows[i]] == [t.as_tuple() for t in wrows[i]], f"sim {i} parts {parts}"
    check(sims[1::2], cv.ENGINE_PERSISTENT)


@pytest.mark.parametrize("threads", [256, 1024])
def test_more_columns_than_threads(threads):
    # grids with more (x-word, y) columns than block threads run their items
    # in rounds over several z-chunks per column; whole and split over 2 and 5
    # blocks, every result matches the oracle
    if threads == 256:
        dims = [(100, 90, 20), (70, 130, 33), (129, 80, 17)]
    else:
        dims = [(200, 180, 12), (161, 250, 9)]
    sims = random_sims(2 * len(dims), 9, dims, ticks=8)
    for s in sims[::2]:
        s.sim_flags |= cv.SIM_NO_EARLY_EXIT
    for parts in (1, 2, 5):
        print(f"threads={threads} max_parts={parts}", flush=True)
        res, rows, _ = cv.batch_run(sims, per_tick=True, block_threads=threads, max_parts=parts)
        for i, s in enumerate(sims):
            info, orows, final = oracle_run(s)
            X, Y, Z = s.dims
            assert res[i].ticks_run == len(orows), f"sim {i} parts {parts}"
            assert [t.as_tuple() for t in rows[i]] == [t.as_tuple() for t in orows], f"sim {i} parts {parts}"
            assert res[i].final_hash == O.canonical_hash(final, X, Y, Z), f"sim {i} parts {parts}"


def test_split_simulations_without_rows():
    # without per-tick rows, split jobs take the one-word per-tick rendezvous
    # (surviving / born totals only; dying totals and the bounding box are
    # combined after the last tick): every result field, the last row and the
    # final cells equal one block per grid and the oracle, with and without
    # early death
    dims = [(70, 40, 30), (100, 33, 21), (65, 64, 64), (96, 50, 37), (128, 41, 77)]
    sims = random_sims(10, 6, dims, ticks=14)
    for s in sims[::2]:
        s.sim_flags |= cv.SIM_NO_EARLY_EXIT

    def key(r):
        return (r.info.totalNumberOfCellsThatDied, r.info.totalNumberOfCellsThatWereBorn,
                r.info.numberOfSimulationTicksSurvived, r.info.status, r.info.maxAliveCount, r.info.finalAliveCount,
                r.classification, r.ticks_run, r.last.as_tuple(), r.final_hash)

    print("whole", flush=True)
    whole, _, _ = cv.batch_run(sims, max_parts=1)
    for parts in (2, 3, 7, 16):
        print(f"max_parts={parts}", flush=True)
        split, _, _ = cv.batch_run(sims, max_parts=parts)
        for i in range(len(sims)):
            assert key(split[i]) == key(whole[i]), f"sim {i} parts {parts}"
    check(sims, cv.ENGINE_PERSISTENT, per_tick=False)


def test_autotuned_uneven_batch():
    # the autotuner picks block size, items per thread and splits; results
    # equal the fixed single-block layout and the oracle
    from cave_amd import workloads as W

    specs = W.search_uneven(96, seed=9)
    specs["ticks"] = np.minimum(specs["ticks"], 24)
    specs["sim_flags"] = cv.SIM_FINAL_HASH
    print("tuned", flush=True)
    res, _, tm, threads = cv.batch_run_np(specs, timing=True)
    assert threads in (256, 512, 1024)
    rep = cv.tuning_report()
    assert rep and {"edge_bucket", "flags", "threads", "items_per_thread", "cell_updates_per_s"} <= set(rep[0])
    assert (tm["end_us"] >= tm["start_us"]).all() and (tm["blocks"] >= 1).all()
    print("fixed", flush=True)
    ref, _ = cv.batch_run_np(specs, block_threads=1024, max_parts=1)
    for f in ("died", "born", "ticks_survived", "status", "max_alive", "final_alive", "classification",
              "ticks_run", "last", "final_hash"):
        assert np.array_equal(res[f], ref[f]), f
    cells = specs["dims"].astype(np.int64).prod(1)
    for i in np.argsort(cells)[:12]:
        sp = specs[i]
        X, Y, Z = (int(v) for v in sp["dims"])
        init = O.init_cells(X, Y, Z, int(sp["M"]), tuple(int(v) for v in sp["spawn"]), int(sp["spawn_mode"]),
                            float(sp["density"]), int(sp["seed"]))
        t, info, _, final = O.run_rule(X, Y, Z, int(sp["M"]), int(sp["flags"]), int(sp["birth"]),
                                       int(sp["survival"]), int(sp["ticks"]), init)
        assert res["ticks_run"][i] == t
        assert res["final_alive"][i] == info.finalAliveCount
        assert res["final_hash"][i] == O.canonical_hash(final, X, Y, Z)


def test_list_calibration_and_block_sizes_agree():
    # 49^3 search grids: the tuner measures this bucket on the list's own
    # simulations (the report's flags carry 64), and every block size -- 256
    # threads with the grids in the arena, 1024 with them in LDS -- gives the
    # same results, which equal the oracle's
    from cave_amd import workloads as W

    specs = W.search_uniform(24, 8)
    specs["sim_flags"] = cv.SIM_FINAL_HASH
    print("tuned", flush=True)
    res, _, _, threads = cv.batch_run_np(specs, timing=True)
    assert threads in (256, 512, 1024)
    assert any(e["flags"] & 64 for e in cv.tuning_report())
    fields = ("died", "born", "ticks_survived", "status", "max_alive", "final_alive", "classification", "ticks_run",
              "last", "final_hash")
    for T in (256, 512, 1024):
        print(f"threads {T}", flush=True)
        got, _ = cv.batch_run_np(specs, block_threads=T)
        for f in fields:
            assert np.array_equal(got[f], res[f]), (T, f)
    for i in range(0, len(specs), 37):
        sp = specs[i]
        X, Y, Z = (int(v) for v in sp["dims"])
        init = O.init_cells(X, Y, Z, int(sp["M"]), tuple(int(v) for v in sp["spawn"]), int(sp["spawn_mode"]),
                            float(sp["density"]), int(sp["seed"]))
        t, info, _, final = O.run_rule(X, Y, Z, int(sp["M"]), int(sp["flags"]), int(sp["birth"]),
                                       int(sp["survival"]), int(sp["ticks"]), init)
        assert res["ticks_run"][i] == t and res["final_alive"][i] == info.finalAliveCount
        assert res["final_hash"][i] == O.canonical_hash(final, X, Y, Z)


@pytest.mark.parametrize("engine", [cv.ENGINE_PERSISTENT, cv.ENGINE_PLAIN])
def test_empty_birth_or_survival_sets(engine):
    # empty sets are ordinary rules here (M = 1 with no birth count packs to
    # the reference's all-zero exhausted-slot word, but a batch spec is a rule)
    sims = [cv.Sim(dims=(20, 17, 9), M=M, flags=fl, ticks=15, birth=b, survival=s, spawn=(10, 9, 5),
                   spawn_mode=cv.SPAWN_RANDOM, density=0.4, seed=7, sim_flags=cv.SIM_FINAL_HASH)
            for M, fl, b, s in [(1, 1, 0, 0), (1, 1, 0, 0x3F), (3, 3, 0, 0), (2, 7, 0x3, 0), (5, 1, 0, 0x6)]]
    check(sims, engine)


@pytest.mark.parametrize("engine", [cv.ENGINE_PERSISTENT, cv.ENGINE_PLAIN])
def test_fev_top_counts(engine):
    # nearly full grids: interior live cells see 26 live neighbours and most
    # dead cells 26 as well; rules that fire only at counts 25 / 26 (mask bit
    # b = count b+1), on LDS-tier (40x20x12, 49^3) and global-tier grids
    sims = []
    for i, dims in enumerate([(40, 20, 12), (49, 49, 49), (70, 40, 30)]):
        for birth, surv in [(1 << 25, 1 << 25), (3 << 24, 3 << 24), (0x3FFFFFF, 1 << 25)]:
            sims.append(cv.Sim(dims=dims, M=1 + 7 * i, flags=7, ticks=6, birth=birth, survival=surv, spawn=dims,
                               spawn_mode=cv.SPAWN_RANDOM, density=0.97, seed=11 + i,
                               sim_flags=cv.SIM_FINAL_HASH | cv.SIM_NO_EARLY_EXIT))
    check(sims, engine)
