"""The persistent engine's active region (exact compute-skip).

By default a tick computes only the cells within one cell of the union of
the live cells' boxes so far (two cells for split jobs that learn the job's
box one tick late); CAVE_SIM_FULL_GRID computes every cell.  Both must give
the same results: GridInfo, classification, every per-tick row (counters and
box) and the final cells.  The cases stress what the region has to cover:
small spawns in large grids (the region grows tick by tick), M = 1 and 2
(cells die right after they stop being live, so live bits left in the
ping-pong live plane two ticks earlier must be rewritten), patterns that die
down or drift, legacy death and no early exit (ticks after the last live
cell), both storage tiers, split jobs on both rendezvous forms, the ERD
shape, and the oracle on a subset."""
import numpy as np
import pytest

import cave_amd as cv
import oracle as O

from test_gpu_batch import check

pytestmark = pytest.mark.gpu


def sims_small_spawns(n, seed, dims_choices, ticks, shape_erd=False, legacy=True):
    rng = np.random.default_rng(seed)
    sims = []
    for i in range(n):
        dims = tuple(int(v) for v in dims_choices[i % len(dims_choices)])
        if shape_erd:
            flags = int((1, 2, 3)[i % 3])
            N = O.lib().orc_erd_neighbor_count(flags)
        else:
            flags = int((1, 2, 3, 4, 5, 6, 7)[i % 7])
            N = O.lib().orc_neighbor_count(flags)
        M = int((1, 2, 3, 5, 9, 15)[i % 6])
        p = (0.08, 0.15, 0.25)[i % 3]
        birth = sum(1 << b for b in range(N) if rng.random() < p) or (1 << int(rng.integers(0, 3)))
        surv = sum(1 << b for b in range(N) if rng.random() < p) or 2
        edge = int(rng.integers(1, max(2, min(dims) // 3)))
        sim_flags = cv.SIM_FINAL_HASH
        if i % 4 == 1:
            sim_flags |= cv.SIM_NO_EARLY_EXIT
        if i % 4 == 2 and legacy:
            sim_flags |= cv.SIM_LEGACY_DEATH
        if shape_erd:
            sim_flags |= cv.SIM_SHAPE_ERD
        sims.append(cv.Sim(dims=dims, M=M, flags=flags, ticks=ticks, birth=birth, survival=surv,
                           spawn=(edge, edge, edge), spawn_mode=cv.SPAWN_RANDOM, density=0.45,
                           seed=int(rng.integers(0, 2**31)), sim_flags=sim_flags))
    return sims


def key(r):
    return (r.info.totalNumberOfCellsThatDied, r.info.totalNumberOfCellsThatWereBorn,
            r.info.numberOfSimulationTicksSurvived, r.info.status, r.info.maxAliveCount, r.info.finalAliveCount,
            r.classification, r.ticks_run, r.last.as_tuple(), r.final_hash)


def full_grid(sims):
    out = []
    for s in sims:
        c = cv.Sim(**{k: getattr(s, k) for k in ("dims", "M", "flags", "ticks", "birth", "survival", "spawn",
                                                  "spawn_mode", "density", "seed", "sim_flags")})
        c.sim_flags |= cv.SIM_FULL_GRID
        out.append(c)
    return out


def compare(sims, per_tick, max_parts=0, label=""):
    print(f"{label}: {len(sims)} sims per_tick={per_tick} max_parts={max_parts}", flush=True)
    a, arows, _ = cv.batch_run(sims, per_tick=per_tick, max_parts=max_parts)
    b, brows, _ = cv.batch_run(full_grid(sims), per_tick=per_tick, max_parts=max_parts)
    for i in range(len(sims)):
        assert key(a[i]) == key(b[i]), f"{label} sim {i} {sims[i]}"
        if per_tick:
            assert [t.as_tuple() for t in arows[i]] == [t.as_tuple() for t in brows[i]], f"{label} sim {i}"
    return a


def test_lds_tier_region_equals_full_grid():
    dims = [(49, 49, 49), (40, 33, 20), (64, 17, 30), (31, 31, 31)]
    sims = sims_small_spawns(36, 1, dims, ticks=30)
    compare(sims, per_tick=True, label="lds")
    compare(sims, per_tick=False, label="lds")


def test_global_tier_region_equals_full_grid():
    dims = [(96, 70, 50), (130, 40, 33), (70, 90, 64)]
    sims = sims_small_spawns(18, 2, dims, ticks=24)
    compare(sims, per_tick=True, max_parts=1, label="global")


@pytest.mark.parametrize("parts", [3, 8])
def test_split_region_equals_full_grid(parts):
    # with per-tick rows split jobs take the general rendezvous (box of this
    # tick), without them the one-word rendezvous (box one tick late)
    dims = [(96, 70, 50), (130, 40, 33), (70, 90, 64)]
    sims = sims_small_spawns(12, 3 + parts, dims, ticks=24)
    compare(sims, per_tick=True, max_parts=parts, label=f"split{parts}")
    compare(sims, per_tick=False, max_parts=parts, label=f"split{parts}")


def test_erd_region_equals_full_grid():
    dims = [(49, 49, 49), (40, 33, 20), (96, 50, 40)]
    sims = sims_small_spawns(18, 4, dims, ticks=24, shape_erd=True)
    compare(sims, per_tick=True, label="erd")
    compare(sims, per_tick=False, max_parts=4, label="erd split")


def test_region_against_the_oracle():
    dims = [(49, 49, 49), (40, 33, 20), (70, 45, 38)]
    sims = sims_small_spawns(14, 5, dims, ticks=25, legacy=False)
    check(sims, cv.ENGINE_PERSISTENT)


def test_gliding_pattern_leaves_no_trail():
    # a sparse rule (birth on 2 or 3 face neighbours, survival on 1) from a
    # tiny spawn, whose live cells spread and vacate cells: every row equals
    # the oracle's, with M = 1 (no decay trail: the live plane written two
    # ticks before must be rewritten), 2 and 4
    sims = []
    for M in (1, 2, 4):
        for seed in range(3):
            sims.append(cv.Sim(dims=(60, 45, 40), M=M, flags=1, ticks=40, birth=0b110, survival=0b1,
                               spawn=(3, 3, 3), spawn_mode=cv.SPAWN_RANDOM, density=0.6, seed=seed,
                               sim_flags=cv.SIM_FINAL_HASH | cv.SIM_NO_EARLY_EXIT))
    check(sims, cv.ENGINE_PERSISTENT)
    compare(sims, per_tick=True, label="glide")
    compare(sims, per_tick=False, max_parts=4, label="glide split")


def test_full_grid_against_the_oracle():
    # the whole-grid flag is a product path too: spot-check it against the oracle
    sims = full_grid(sims_small_spawns(7, 6, [(33, 20, 17)], ticks=12, legacy=False))
    check(sims, cv.ENGINE_PERSISTENT)
