"""Continuous-batching server (cave_serve_*): every served result equals the
same request alone, the batched engine and the CPU oracle; admission keeps
resident state within the HBM budget; split requests and kernel relaunches
after an idle close give the same results."""
import time

import numpy as np
import pytest

import cave_amd as cv
import oracle as O

pytestmark = pytest.mark.gpu


def requests(n, seed, edges=(12, 20, 33, 48, 64, 90), ticks=(5, 40)):
    rng = np.random.default_rng(seed)
    sims = []
    for i in range(n):
        e = int(edges[i % len(edges)])
        dims = (e, max(3, e - int(rng.integers(0, 6))), max(3, e + int(rng.integers(-4, 5))))
        flags = int((1, 3, 7, 5)[i % 4])
        N = O.lib().orc_neighbor_count(flags)
        birth = sum(1 << b for b in range(N) if rng.random() < 0.15)
        surv = sum(1 << b for b in range(N) if rng.random() < 0.2)
        sims.append(cv.Sim(dims=dims, M=int(rng.integers(1, 16)), flags=flags,
                           ticks=int(rng.integers(ticks[0], ticks[1] + 1)), birth=birth, survival=surv,
                           spawn=tuple(max(1, d // 2) for d in dims), spawn_mode=cv.SPAWN_RANDOM, density=0.25,
                           seed=int(rng.integers(0, 2**31)), sim_flags=cv.SIM_FINAL_HASH))
    return sims


def key(r):
    return (r.info.totalNumberOfCellsThatDied, r.info.totalNumberOfCellsThatWereBorn,
            r.info.numberOfSimulationTicksSurvived, r.info.status, r.info.maxAliveCount, r.info.finalAliveCount,
            r.classification, r.ticks_run, r.last.as_tuple(), r.final_hash)


def serve_all(srv, sims, gap_s=0.0):
    for i, s in enumerate(sims):
        srv.submit(s, i)
        if gap_s:
            time.sleep(gap_s)
    done = srv.drain()
    assert sorted(c.request_id for c in done) == list(range(len(sims)))
    assert all(c.status == 0 for c in done)
    return {c.request_id: c for c in done}


def test_served_equals_alone_batch_and_oracle():
    sims = requests(40, 1)
    print("mixed", flush=True)
    with cv.Server() as srv:
        mixed = serve_all(srv, sims)
        print("alone", flush=True)
        alone = {}
        for i, s in enumerate(sims):
            srv.submit(s, 1000 + i)
            (c,) = srv.drain()
            alone[i] = c
        st = srv.stats()
    assert st["admitted"] == st["finished"] == 80 and st["resident_bytes"] == 0
    for i in range(len(sims)):
        assert key(mixed[i].result) == key(alone[i].result), f"request {i}"
        assert mixed[i].admit_s >= mixed[i].submit_s and mixed[i].finish_s >= mixed[i].admit_s
    print("batch", flush=True)
    res, _, _ = cv.batch_run(sims)
    for i, r in enumerate(res):
        assert key(mixed[i].result) == key(r), f"request {i}"
    for i, s in enumerate(sims):
        if np.prod(s.dims) * s.ticks > 3e6:
            continue
        X, Y, Z = s.dims
        init = O.init_cells(X, Y, Z, s.M, s.spawn, s.spawn_mode, s.density, s.seed)
        t, info, _, final = O.run_rule(X, Y, Z, s.M, s.flags, s.birth, s.survival, s.ticks, init)
        r = mixed[i].result
        assert (r.ticks_run, r.info.finalAliveCount, r.info.totalNumberOfCellsThatDied) == (
            t, info.finalAliveCount, info.totalNumberOfCellsThatDied), f"request {i}"
        assert r.final_hash == O.canonical_hash(final, X, Y, Z), f"request {i}"


def test_interleaved_arrivals_and_split_requests():
    # a small split target splits the larger grids over several blocks while
    # other requests keep arriving
    sims = requests(36, 2, edges=(40, 72, 96, 130), ticks=(10, 30))
    ref, _, _ = cv.batch_run(sims)
    with cv.Server(split_target_us=200) as srv:
        got = serve_all(srv, sims, gap_s=0.0005)
    assert any(c.blocks > 1 for c in got.values())
    for i, r in enumerate(ref):
        assert key(got[i].result) == key(r), f"request {i} blocks {got[i].blocks}"


def test_admission_budget():
    # a 100^3 request holds ~1.0 MB of resident state plus 0.5 MB of final-cell
    # staging: the budget admits about two at a time
    sims = requests(12, 3, edges=(100,), ticks=(20, 30))
    budget = 3 * 100 * 100 * 100
    ref, _, _ = cv.batch_run(sims)
    with cv.Server(hbm_budget_bytes=budget) as srv:
        got = serve_all(srv, sims)
        st = srv.stats()
        assert 0 < st["peak_resident_bytes"] <= budget
        big = cv.Sim(dims=(400, 400, 400), M=2, flags=1, ticks=2, birth=1, survival=1, spawn=(4, 4, 4))
        with pytest.raises(cv.CaveError):
            srv.submit(big, 99)
    for i, r in enumerate(ref):
        assert key(got[i].result) == key(r), f"request {i}"


def test_relaunch_after_idle_close():
    sims = requests(10, 4, edges=(24, 70), ticks=(5, 12))
    ref, _, _ = cv.batch_run(sims)
    with cv.Server(idle_exit_us=2000) as srv:
        first = serve_all(srv, sims[:5])
        time.sleep(0.05)  # the resident kernel closes after 2 ms without claims
        srv2 = {}
        for i, s in enumerate(sims[5:], start=5):
            srv.submit(s, i)
        for c in srv.drain():
            srv2[c.request_id] = c
    got = {**first, **srv2}
    for i, r in enumerate(ref):
        assert key(got[i].result) == key(r), f"request {i}"


def test_serving_envelope_edge_192_500_ticks_against_the_cpu():
    """The serving workload's own envelope: the 24 costliest requests of a 2 s
    stream of it (edges 176..191, F/FE/FEV, C in 2..16, 374..498-tick budgets,
    about half of them alive to the end), served together so the large ones
    split over many blocks, each compared with the fast CPU form (ticks, every
    grid-info field, final-cell hash)."""
    from cave_amd import workloads as W
    from cpu_ref import cpu_threads, fast_sims

    _, specs = W.serving_requests(2000.0, 2.0)
    cost = specs["dims"][:, 0].astype(np.int64) ** 3 * specs["ticks"]
    idx = np.argsort(-cost, kind="stable")[:24]
    sub = np.ascontiguousarray(specs[idx])
    sub["sim_flags"] |= cv.SIM_FINAL_HASH
    cticks, cinfo, ran, _, _, chash = O.run_sims_fast(fast_sims(sub), n_threads=cpu_threads(), hashes=True)
    assert ran == len(sub) and int(cticks.max()) > 400
    print(f"serving {len(sub)} envelope requests", flush=True)
    with cv.Server() as srv:
        for k in range(len(sub)):
            srv.submit_row(sub, k, k)
        got = {c.request_id: c for c in srv.drain()}
    assert sorted(got) == list(range(len(sub)))
    assert sum(c.blocks > 1 for c in got.values()) >= len(sub) // 2
    for k in range(len(sub)):
        c = got[k]
        r = c.result
        assert c.status == 0, f"request {k}"
        assert r.ticks_run == cticks[k], f"request {k}"
        for f, _ in O.GridInfo._fields_:
            assert getattr(r.info, f) == getattr(cinfo[k], f), (k, f)
        assert r.final_hash == int(chash[k]), f"request {k} blocks {c.blocks}"
