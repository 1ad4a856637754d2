"""Single-grid step on the GPU vs the CPU oracle: bit-exact, state by state and
tick by tick, through the C-ABI (cave_step_*)."""
import json
import os

import numpy as np
import pytest

import cave_amd as cv
import oracle as O

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
KERNELS = [cv.KERNEL_GATHER, cv.KERNEL_STREAM, cv.KERNEL_STREAM_NIBBLE]
KNAME = {cv.KERNEL_GATHER: "gather", cv.KERNEL_STREAM: "stream", cv.KERNEL_STREAM_NIBBLE: "stream_nibble"}


def stats_row(st):
    d, dy, b, s, live, ch, bmin, bmax = st.as_tuple()
    return {"dead": d, "dying": dy, "born": b, "surviving": s, "live": live, "changed": ch,
            "bbox_min": list(bmin), "bbox_max": list(bmax)}


def fixture_check(case, kernel):
    sp = case["spec"]
    X, Y, Z = sp["dims"]
    print(f"case dims={sp['dims']} M={sp['M']} flags={sp['flags']} kernel={KNAME[kernel]}", flush=True)
    with cv.Step((X, Y, Z), sp["M"], sp["flags"], sp["birth"], sp["survival"], spawn=sp["spawn"],
                 spawn_mode=sp["spawn_mode"], density=sp["density"], seed=sp["seed"], kernel=kernel) as h:
        assert h.kernel == kernel
        rows = case["rows"]
        assert O.canonical_hash(h.cells(), X, Y, Z) == rows[0]["hash"]
        for r in rows[1:]:
            (st,), _ = h.run(1)
            got = stats_row(st)
            want = {k: r[k] for k in got}
            assert got == want, f"tick {r['tick']}"
            assert O.canonical_hash(h.cells(), X, Y, Z) == r["hash"], f"tick {r['tick']}"


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: KNAME[k])
def test_fixture_49_FE_C2(kernel):
    with open(os.path.join(HERE, "golden", "step_49_FE_C2.json")) as f:
        fixture_check(json.load(f), kernel)


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: KNAME[k])
def test_fixtures_small(kernel):
    with open(os.path.join(HERE, "golden", "step_small.json")) as f:
        for case in json.load(f):
            fixture_check(case, kernel)


RANDOM_CASES = [
    # (X, Y, Z, M, flags)
    (1, 1, 1, 1, 7), (2, 3, 4, 2, 1), (7, 3, 2, 3, 3), (33, 5, 4, 1, 7), (31, 6, 9, 15, 2),
    (64, 9, 7, 4, 4), (100, 17, 13, 7, 7), (65, 2, 30, 5, 5), (129, 11, 6, 2, 6), (96, 33, 5, 1, 3),
    # rows of more than 256 cells fill a whole 16-lane DPP row: the full-row
    # variant of the streaming kernel, every neighbourhood
    (300, 9, 6, 3, 1), (300, 9, 6, 4, 2), (300, 9, 6, 5, 3), (300, 9, 6, 6, 4), (300, 9, 6, 2, 5),
    (300, 9, 6, 15, 6), (300, 9, 6, 7, 7), (257, 5, 4, 1, 7), (511, 4, 3, 9, 3), (512, 6, 5, 7, 7),
    # rows of 513..1024 cells take 32 lanes (two x-rows per wave) and rows of
    # 1025..2048 a whole wave: x-neighbours cross 16-lane rows with wave shifts
    (513, 9, 4, 3, 7), (600, 9, 5, 4, 3), (777, 10, 3, 2, 5), (1024, 9, 4, 6, 7), (1025, 5, 4, 1, 1),
    (1100, 5, 6, 7, 7), (1500, 7, 3, 15, 2), (2047, 5, 4, 5, 6), (2048, 5, 3, 3, 7), (2048, 3, 17, 2, 4),
]


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: KNAME[k])
@pytest.mark.parametrize("X,Y,Z,M,flags", RANDOM_CASES)
def test_random_grids_state_by_state(kernel, X, Y, Z, M, flags):
    rng = np.random.default_rng(X * 7919 + Y * 31 + Z + M * 1000 + flags)
    n = O.lib().orc_neighbor_count(flags)
    birth = int(rng.integers(1, 1 << n))
    surv = int(rng.integers(1, 1 << n))
    cells = rng.integers(0, M + 1, size=O.cell_words(X, Y, Z) * 8).astype(np.uint32)
    # build a valid packed state (padding nibbles zero) from random nibbles
    dense = cells.reshape(-1)[: X * Y * Z].astype(np.uint8).reshape(Z, Y, X)
    import np_model as NP
    packed = NP.pack(dense)
    w0 = ((M - 1) & 0xF) << 60 | birth
    w1 = flags << 60 | surv
    print(f"case {X}x{Y}x{Z} M={M} flags={flags} kernel={KNAME[kernel]}", flush=True)
    with cv.Step((X, Y, Z), M, flags, birth, surv, initial_cells=packed, kernel=kernel) as h:
        ref = packed
        for t in range(8):
            ref, st_ref = O.step(ref, X, Y, Z, w0, w1)
            (st,), _ = h.run(1)
            assert st.as_tuple() == st_ref.as_tuple(), f"tick {t}"
            assert np.array_equal(h.cells(), ref), f"tick {t}"


WRAP_CASES = [
    # (X, Y, Z, M, flags): tiny tori count the same cell once per offset that reaches it
    (1, 1, 1, 1, 7), (2, 3, 1, 2, 3), (3, 2, 2, 4, 5), (7, 5, 3, 1, 7), (9, 8, 10, 3, 1),
    (33, 17, 9, 2, 3), (64, 3, 5, 15, 6), (100, 31, 7, 5, 7), (17, 1, 40, 1, 2),
    # the streaming kernel's forms: several y tiles and z tiles, rows of 257..512
    # cells (one DPP row), 513..1024 (half a wave), 1025..2048 (a wave), a
    # whole last group, and rows per wave that do not divide Y
    (33, 70, 9, 3, 7), (40, 9, 70, 2, 3), (300, 9, 5, 4, 7), (512, 5, 3, 7, 7), (600, 5, 4, 2, 3),
    (1100, 3, 3, 3, 5), (2048, 2, 2, 1, 7),
]
WRAP_KERNELS = [cv.KERNEL_AUTO, cv.KERNEL_GATHER]


@pytest.mark.parametrize("kernel", WRAP_KERNELS, ids=["auto", "gather"])
@pytest.mark.parametrize("X,Y,Z,M,flags", WRAP_CASES)
def test_wrap_grids_state_by_state(X, Y, Z, M, flags, kernel):
    """The W flag wraps every axis (the render path's step,
    rayMarchCompute.slang:80-108): GPU vs oracle, state and counters per tick,
    on the planar streaming kernel (the default for W) and the gather kernel."""
    import np_model as NP

    rng = np.random.default_rng(X * 131 + Y * 17 + Z + M * 999 + flags)
    n = O.lib().orc_neighbor_count(flags)
    birth = int(rng.integers(1, 1 << n))
    surv = int(rng.integers(1, 1 << n))
    dense = rng.integers(0, M + 1, size=(Z, Y, X)).astype(np.uint8)
    packed = NP.pack(dense)
    w0 = ((M - 1) & 0xF) << 60 | birth
    w1 = (flags | cv.WRAP) << 60 | surv
    print(f"wrap case {X}x{Y}x{Z} M={M} flags={flags} kernel={kernel}", flush=True)
    with cv.Step((X, Y, Z), M, flags | cv.WRAP, birth, surv, initial_cells=packed, kernel=kernel) as h:
        assert h.kernel == (cv.KERNEL_GATHER if kernel == cv.KERNEL_GATHER else cv.KERNEL_STREAM)
        ref = packed
        for t in range(6):
            ref, st_ref = O.step_render(ref, X, Y, Z, w0, w1)
            (st,), _ = h.run(1)
            assert st.as_tuple() == st_ref.as_tuple(), f"tick {t}"
            assert np.array_equal(h.cells(), ref), f"tick {t}"


def test_wrap_is_not_on_the_nibble_layout():
    with cv.Step((8, 8, 8), 1, cv.FACES | cv.WRAP, 1, 1, kernel=cv.KERNEL_STREAM) as h:
        assert h.kernel == cv.KERNEL_STREAM
    with pytest.raises(cv.CaveError) as e:
        cv.Step((8, 8, 8), 1, cv.FACES | cv.WRAP, 1, 1, kernel=cv.KERNEL_STREAM_NIBBLE)
    assert e.value.code == cv.CAVE_E_UNSUPPORTED


@pytest.mark.parametrize("density", [0.35, 0.2, 1.0 / 3.0, 1.0])
def test_render_random_spawn(density):
    """cellInit.slang's initial cells: unsalted hash, float threshold."""
    X, Y, Z, M, spawn = 70, 33, 21, 6, (40, 20, 15)
    with cv.Step((X, Y, Z), M, cv.FACES | cv.EDGES, 1, 1, spawn=spawn, spawn_mode=cv.SPAWN_RENDER_RANDOM,
                 density=density, seed=12345) as h:
        assert np.array_equal(h.cells(), O.init_cells_render(X, Y, Z, M, spawn, 0, density))


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: KNAME[k])
def test_multi_tick_run_matches_single_ticks(kernel):
    X, Y, Z, M, flags = 96, 40, 24, 7, 7
    rng = cv.Rng(5)
    birth, surv = rng.rule(26, 0.15), rng.rule(26, 0.15)
    a = cv.Step((X, Y, Z), M, flags, birth, surv, spawn_mode=cv.SPAWN_RANDOM, density=0.2, seed=9, kernel=kernel)
    b = cv.Step((X, Y, Z), M, flags, birth, surv, spawn_mode=cv.SPAWN_RANDOM, density=0.2, seed=9,
                kernel=cv.KERNEL_GATHER)
    rows_a, _ = a.run(30)
    rows_b = []
    for _ in range(30):
        (st,), _ = b.run(1)
        rows_b.append(st)
    assert [r.as_tuple() for r in rows_a] == [r.as_tuple() for r in rows_b]
    assert np.array_equal(a.cells(), b.cells())
    a.close()
    b.close()


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: KNAME[k])
@pytest.mark.parametrize("M", [1, 7, 15])
def test_fev_top_counts(kernel, M):
    """FEV counts up to 26: a fully live grid with isolated dead cells, so
    interior live cells see 26 live neighbours and most dead cells 26 too,
    under rules that fire only at the top counts (mask bit b = count b+1)."""
    import np_model as NP

    X, Y, Z = 70, 12, 10
    rng = np.random.default_rng(M)
    dense = np.full((Z, Y, X), M, dtype=np.uint8)
    dense[rng.random((Z, Y, X)) < 0.03] = 0
    packed = NP.pack(dense)
    for birth, surv in [(1 << 25, 1 << 25), (3 << 24, 3 << 24), (0x3FFFFFF, 1 << 25)]:
        w0 = ((M - 1) & 0xF) << 60 | birth
        w1 = 7 << 60 | surv
        print(f"FEV top counts M={M} B={birth:#x} S={surv:#x} kernel={KNAME[kernel]}", flush=True)
        with cv.Step((X, Y, Z), M, 7, birth, surv, initial_cells=packed, kernel=kernel) as h:
            ref = packed
            for t in range(4):
                ref, st_ref = O.step(ref, X, Y, Z, w0, w1)
                (st,), _ = h.run(1)
                assert st.as_tuple() == st_ref.as_tuple(), f"tick {t}"
                assert np.array_equal(h.cells(), ref), f"tick {t}"


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: KNAME[k])
def test_garbage_row_padding_is_ignored(kernel):
    # the reference never reads padding nibbles as neighbours; live-valued
    # garbage in them must not change any result
    X, Y, Z, M, flags = 37, 6, 5, 3, 7
    rng = np.random.default_rng(77)
    import np_model as NP
    dense = rng.integers(0, M + 1, size=(Z, Y, X)).astype(np.uint8)
    packed = NP.pack(dense)
    dirty = packed.copy().reshape(Z, Y, -1)
    dirty[:, :, -1] |= np.uint32(0x33300000)  # last word: nibbles 5..7 (x = 37..39, padding) set to M
    dirty = dirty.reshape(-1)
    birth, surv = 0b1110, 0b111000
    w0 = ((M - 1) & 0xF) << 60 | birth
    w1 = flags << 60 | surv
    with cv.Step((X, Y, Z), M, flags, birth, surv, initial_cells=dirty, kernel=kernel) as h:
        ref = packed
        for t in range(3):
            ref, st_ref = O.step(ref, X, Y, Z, w0, w1)
            (st,), _ = h.run(1)
            assert st.as_tuple() == st_ref.as_tuple(), f"tick {t}"
            assert np.array_equal(h.cells(), ref), f"tick {t}"


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: KNAME[k])
def test_graph_replays_match_eager_and_oracle(kernel):
    """A run of ticks is a captured HIP graph replayed per call: replays of
    both buffer parities, interleaved with eager work and host syncs, equal
    one-by-one launches and the oracle after every call."""
    import os as _os

    X, Y, Z, M, flags = 70, 23, 19, 4, 7
    rng = np.random.default_rng(5)
    n = O.lib().orc_neighbor_count(flags)
    birth, surv = int(rng.integers(1, 1 << n)), int(rng.integers(1, 1 << n))
    import np_model as NP
    packed = NP.pack(rng.integers(0, M + 1, size=(Z, Y, X)).astype(np.uint8))
    w0 = ((M - 1) & 0xF) << 60 | birth
    w1 = flags << 60 | surv
    eager = cv.Step((X, Y, Z), M, flags, birth, surv, initial_cells=packed, kernel=kernel)
    _os.environ["CAVE_STEP_GRAPHS"] = "1"
    try:
        graphed = cv.Step((X, Y, Z), M, flags, birth, surv, initial_cells=packed, kernel=kernel)
    finally:
        del _os.environ["CAVE_STEP_GRAPHS"]
    ref = packed
    with graphed as h, eager:
        for ticks in (5, 5, 3, 5, 1, 5, 3, 3):
            rows, _ = h.run(ticks)
            erows, _ = eager.run(ticks)
            for t in range(ticks):
                ref, st = O.step(ref, X, Y, Z, w0, w1)
                assert rows[t].as_tuple() == st.as_tuple() == erows[t].as_tuple()
            assert np.array_equal(h.cells(), ref) and np.array_equal(eager.cells(), ref)


@pytest.mark.parametrize("kernel", KERNELS, ids=lambda k: KNAME[k])
def test_rows_after_runs_without_rows_and_across_chunks(kernel):
    """Per-tick rows come back through host-mapped memory and a completion
    flag; runs without rows skip the reduction.  Interleaved runs with and
    without rows, and a run longer than one chunk of ticks (1024 on a grid
    this small: two reductions, two flags in one call), against the oracle
    tick by tick (searchSimulationCube.slang:78-247)."""
    X, Y, Z, M, flags = 21, 9, 11, 3, 3
    rng = cv.Rng(3)  # a rule whose cells still change on every tick past 1,000
    birth, surv = rng.rule(18, 0.3), rng.rule(18, 0.4)
    w0, w1 = ((M - 1) << 60) | birth, (flags << 60) | surv
    packed = O.init_cells(X, Y, Z, M, (X, Y, Z), 0, 0.3, 5)
    print(f"dims={X}x{Y}x{Z} M={M} kernel={KNAME[kernel]}", flush=True)
    ref = packed
    with cv.Step((X, Y, Z), M, flags, birth, surv, initial_cells=packed, kernel=kernel) as h:
        for ticks, want_rows in ((3, False), (4, True), (2, False), (1, True), (1100, True), (5, False), (3, True)):
            rows, _ = h.run(ticks, stats=want_rows)
            for t in range(ticks):
                ref, st = O.step(ref, X, Y, Z, w0, w1)
                if want_rows:
                    assert rows[t].as_tuple() == st.as_tuple(), f"run of {ticks}, tick {t}"
            assert st.as_tuple()[5] > 0  # still changing
            assert rows is None or len(rows) == ticks
            assert np.array_equal(h.cells(), ref), f"cells after a run of {ticks}"
