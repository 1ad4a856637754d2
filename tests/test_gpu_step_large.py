"""The single-grid step at the sizes the benchmark times, on the GPU vs the
fast CPU form (oracle/cpu_fast.c, itself held bit-exact to the scalar oracle
by tests/test_cpu_fast.py): initial cells, every tick's counters and bounding
box, and the cells after the last tick, through the C-ABI (cave_step_*).

512^3 is the benchmark grid itself: the streaming kernel's default plan there
is 1024 blocks of 4 waves (a multiple of 8, so the XCD block remap is taken),
64 y tiles x 64 z tiles of 8 planes, odd z tiles walking downwards, and rows
that fill a whole 16-lane DPP row.  256^3 takes 128 blocks (also remapped)
with two x-rows per DPP row.  200x130x70 has a partial last group and a
partial last y and z tile.  Reference step: searchSimulationCube.slang:159-247.
"""
import numpy as np
import pytest

import cave_amd as cv
import oracle as O
from cpu_ref import cpu_threads

pytestmark = pytest.mark.gpu

FEV = cv.FACES | cv.EDGES | cv.CORNERS


def bench_rule():
    """The benchmark's rule: splitmix64 seed 2, p = 0.15 per count 1..26."""
    rng = cv.Rng(2)
    return rng.rule(26, 0.15), rng.rule(26, 0.15)


def busy_rule():
    """A rule with births and survivals around the mean live count of a
    Bernoulli(0.2) FEV grid (26 * 0.2 = 5.2), so every bucket stays busy."""
    return cv.rule_string_to_bitmask("4-6,9"), cv.rule_string_to_bitmask("3-7,11")


def compare_run(dims, M, flags, birth, surv, density, seed, kernel, ticks):
    X, Y, Z = dims
    print(f"case {X}x{Y}x{Z} M={M} flags={flags} B={birth:#x} S={surv:#x} kernel={kernel} ticks={ticks}",
          flush=True)
    init = O.init_cells(X, Y, Z, M, dims, 0, density, seed)
    fast = O.FastGrid(init, X, Y, Z)
    with cv.Step(dims, M, flags, birth, surv, spawn=dims, spawn_mode=cv.SPAWN_RANDOM, density=density,
                 seed=seed, kernel=kernel) as h:
        assert h.kernel == kernel
        assert np.array_equal(h.cells(), init), "initial cells"
        del init
        want = fast.run(M, flags, birth, surv, ticks, n_threads=cpu_threads(), stats=True)
        got, _ = h.run(ticks)
        for t in range(ticks):
            assert got[t].as_tuple() == want[t].as_tuple(), f"tick {t}"
        assert np.array_equal(h.cells(), fast.cells()), f"cells after {ticks} ticks"
    fast.close()


def test_bench_grid_512_fev_c8():
    """The benchmark workload itself: 512^3, FEV, M = 7, its rule, Bernoulli(0.2)
    with seed 42; 10 ticks covers the whole dying chain (6 dying states)."""
    b, s = bench_rule()
    compare_run((512, 512, 512), 7, FEV, b, s, 0.2, 42, cv.KERNEL_STREAM, 10)


def test_busy_rule_512_fev_c8():
    b, s = busy_rule()
    compare_run((512, 512, 512), 7, FEV, b, s, 0.2, 7, cv.KERNEL_STREAM, 6)


@pytest.mark.parametrize("kernel", [cv.KERNEL_STREAM, cv.KERNEL_STREAM_NIBBLE], ids=["stream", "stream_nibble"])
@pytest.mark.parametrize("flags", [cv.FACES | cv.EDGES, FEV], ids=["FE", "FEV"])
def test_256_multi_tile(kernel, flags):
    b, s = busy_rule()
    compare_run((256, 256, 256), 3, flags, b & 0x3FFFF, s & 0x3FFFF, 0.25, 11, kernel, 6)


@pytest.mark.parametrize("kernel", [cv.KERNEL_STREAM, cv.KERNEL_STREAM_NIBBLE], ids=["stream", "stream_nibble"])
def test_ragged_multi_tile(kernel):
    b, s = busy_rule()
    compare_run((200, 130, 70), 4, FEV, b, s, 0.3, 5, kernel, 6)


# ---- rows wider than 512 cells, large buffers, more than 2^32 cells ----

@pytest.mark.parametrize("kernel", [cv.KERNEL_STREAM, cv.KERNEL_STREAM_NIBBLE], ids=["stream", "stream_nibble"])
@pytest.mark.parametrize("dims,M,flags", [
    ((600, 96, 80), 5, FEV),                   # 32 lanes per x-row, partial last group
    ((1024, 128, 96), 3, cv.FACES | cv.EDGES),  # 32 lanes per x-row, full
    ((1100, 70, 40), 7, FEV),                  # a whole wave per x-row, partial
    ((2048, 64, 48), 2, cv.FACES),             # a whole wave per x-row, full
], ids=["600", "1024", "1100", "2048"])
def test_wide_rows(kernel, dims, M, flags):
    b, s = busy_rule()
    n = O.lib().orc_neighbor_count(flags)
    mask = (1 << n) - 1
    compare_run(dims, M, flags, b & mask, s & mask, 0.25, 21, kernel, 5)


def compare_words(dims, M, flags, birth, surv, ticks, seed):
    """Random 32-bit words as the initial cells: with M = 15 every nibble is a
    state (0 dead, 1..14 dying, 15 live), and X a multiple of 8 leaves no row
    padding.  The fast CPU form starts from the same words."""
    X, Y, Z = dims
    assert X % 8 == 0
    print(f"case {X}x{Y}x{Z} M={M} flags={flags} random words, ticks={ticks}", flush=True)
    words = np.random.default_rng(seed).integers(0, 1 << 32, size=O.cell_words(X, Y, Z), dtype=np.uint64)
    words = words.astype(np.uint32)
    fast = O.FastGrid(words, X, Y, Z)
    with cv.Step(dims, M, flags, birth, surv, initial_cells=words, kernel=cv.KERNEL_STREAM) as h:
        assert h.kernel == cv.KERNEL_STREAM
        del words
        want = fast.run(M, flags, birth, surv, ticks, n_threads=cpu_threads(), stats=True)
        got, _ = h.run(ticks)
        for t in range(ticks):
            assert got[t].as_tuple() == want[t].as_tuple(), f"tick {t}"
        assert np.array_equal(h.cells(), fast.cells()), f"cells after {ticks} ticks"
    fast.close()


def test_1024_cubed():
    compare_words((1024, 1024, 1024), 15, FEV, cv.rule_string_to_bitmask("1,2"),
                  cv.rule_string_to_bitmask("1-3"), 3, 31)


def test_buffer_beyond_1gib():
    # 1024 x 1024 x 2100: 1.10 GB per state buffer
    compare_words((1024, 1024, 2100), 15, cv.FACES | cv.EDGES, cv.rule_string_to_bitmask("2,3"),
                  cv.rule_string_to_bitmask("1,2"), 2, 32)


def test_more_than_2_32_cells():
    """2048 x 1024 x 2049 = 4.30e9 cells (2.15 GB per buffer): the per-tick
    counters wrap modulo 2^32 on both sides, as the reference's u32
    counters do."""
    compare_words((2048, 1024, 2049), 15, FEV, cv.rule_string_to_bitmask("1"), cv.rule_string_to_bitmask("1,2"),
                  2, 33)


@pytest.mark.parametrize("edge", [950, 1500, 2023])
def test_reference_large_grid_sizes(edge):
    """The reference's large-grid workloads (src/testing/testRunner.cpp:200-229,
    RunSimulationTicks at :21-90): the render path's initial cells (unsalted
    hash, density 0.5) in a centred spawn box of edge/5, birth {4}, survival
    {4}, max cell state 5, F neighbourhood, no wrap, 5 ticks.  Activity
    spreads at most one cell per tick beyond the box, so: the fast CPU form
    reruns a window of the box grown by ticks + 1 cells (cut from the GPU's
    own initial cells, whose init kernel the smaller tests pin against the
    oracle) and must match the GPU's window; every cell outside the window
    must be 0 before and after; and the whole-grid counters must be the
    window's plus the outside cells, dead (modulo 2^32 on 2023^3)."""
    ticks, M, flags = 5, 5, cv.FACES
    birth, surv = cv.rule_string_to_bitmask("4"), cv.rule_string_to_bitmask("4")
    E, S = edge, edge // 5
    lo, hi = (E - S) // 2, (E + S) // 2
    w0, w1 = max(0, lo - ticks - 1), min(E, hi + ticks + 1)
    x0, x1 = w0 // 8 * 8, min(-(-w1 // 8) * 8, -(-E // 8) * 8)
    upr = -(-E // 8)
    WX, WY, WZ = x1 - x0, w1 - w0, w1 - w0
    print(f"reference large-grid case {E}^3 spawn {S}: window x {x0}..{x1}, y/z {w0}..{w1}", flush=True)

    def window(words):
        g = words.reshape(E, E, upr)
        win = np.ascontiguousarray(g[w0:w1, w0:w1, x0 // 8:x1 // 8]).reshape(-1)
        g[w0:w1, w0:w1, x0 // 8:x1 // 8] = 0
        assert not g.any(), "a cell outside the window is not 0"
        return win

    with cv.Step((E, E, E), M, flags, birth, surv, spawn=(S, S, S), spawn_mode=cv.SPAWN_RENDER_RANDOM,
                 density=0.5, kernel=cv.KERNEL_STREAM) as h:
        init = window(h.cells())
        assert init.any()
        fast = O.FastGrid(init, WX, WY, WZ)
        del init
        want = fast.run(M, flags, birth, surv, ticks, n_threads=cpu_threads(), stats=True)
        got, _ = h.run(ticks)
        # cells of the grid outside the window (window columns past X, if any,
        # are dead on the CPU side and absent on the GPU side: the same offset)
        outside = E ** 3 - WX * WY * WZ
        for t in range(ticks):
            d, dy, b, s, live, ch, bmin, bmax = want[t].as_tuple()
            exp = ((d + outside) % 2**32, dy, b, s, live, ch,
                   tuple(v + o for v, o in zip(bmin, (x0, w0, w0))) if live else bmin,
                   tuple(v + o for v, o in zip(bmax, (x0, w0, w0))) if live else bmax)
            assert got[t].as_tuple() == exp, f"tick {t}"
        assert np.array_equal(window(h.cells()), fast.cells()), "cells after 5 ticks"
    fast.close()
