"""The fast CPU form (oracle/cpu_fast.c) against the scalar restatement.

The fast form is the host-CPU baseline that bench.py times; it is only a
baseline if it computes the same thing.  Every check here compares it with
orc_step_stats / orc_run_rule on the same cells: cells state by state and the
per-tick statistics (snapshot counts, live / changed, bounding box) tick by
tick, over every neighbourhood flag set, M in 1..15, row widths on both sides
of the 64-cell word boundary, and 1..5 threads (slab edges inside the grid).
"""
import numpy as np
import pytest

import oracle as O

UINT32_MAX = 0xFFFFFFFF


def _check(X, Y, Z, M, flags, birth, surv, cells, ticks, threads):
    g = O.FastGrid(cells, X, Y, Z)
    try:
        fast = g.run(M, flags, birth, surv, ticks, n_threads=threads)
        c = cells
        for t in range(ticks):
            c, s = O.step(c, X, Y, Z, ((M - 1) << 60) | birth, (flags << 60) | surv)
            assert fast[t].as_tuple() == s.as_tuple(), (X, Y, Z, M, flags, t)
        assert np.array_equal(g.cells(), c), (X, Y, Z, M, flags)
    finally:
        g.close()


@pytest.mark.parametrize("flags", range(1, 8))
def test_every_neighbourhood(flags):
    rng = np.random.default_rng(100 + flags)
    for _ in range(4):
        X, Y, Z = (int(v) for v in rng.integers(3, 40, 3))
        M = int(rng.integers(1, 16))
        birth = int(rng.integers(0, 1 << 26)) & int(rng.integers(0, 1 << 26))
        surv = int(rng.integers(0, 1 << 26)) & int(rng.integers(0, 1 << 26))
        cells = O.init_cells(X, Y, Z, M, (X, Y, Z), 0, float(rng.uniform(0.1, 0.6)), int(rng.integers(1 << 30)))
        _check(X, Y, Z, M, flags, birth, surv, cells, 5, int(rng.integers(1, 4)))


@pytest.mark.parametrize("X", [1, 2, 63, 64, 65, 127, 128, 129, 200])
def test_row_widths(X):
    """Rows that end inside, at and just past a 64-cell word: the x
    neighbours cross words, and padding cells stay 0 and are never born."""
    Y, Z = 9, 7
    # birth on every count: any padding cell with a live neighbour would be born
    cells = O.init_cells(X, Y, Z, 5, (X, Y, Z), 0, 0.3, 7 + X)
    _check(X, Y, Z, 5, 7, (1 << 26) - 1, 0x1F0, cells, 4, 2)


def test_all_states_and_every_count():
    """Cells in every state 0..M, and rules that allow each single count in
    turn (each one-hot count decode of the 5-bit sum)."""
    X, Y, Z, M = 23, 19, 17, 15
    rng = np.random.default_rng(5)
    cells = np.zeros(O.cell_words(X, Y, Z), dtype=np.uint32)
    upr = (X + 7) // 8
    v = rng.integers(0, M + 1, size=(Z, Y, X))
    v[rng.random((Z, Y, X)) < 0.5] = M  # dense live cells reach the high counts
    for z in range(Z):
        for y in range(Y):
            for x in range(X):
                cells[(z * Y + y) * upr + x // 8] |= np.uint32(int(v[z, y, x]) << (4 * (x % 8)))
    for n in range(1, 27):
        _check(X, Y, Z, M, 7, 1 << (n - 1), 1 << (n - 1), cells, 1, 1)


def test_threads_and_thin_slabs():
    """More threads than a slab of a few planes: every slab edge is a z
    neighbour read from another thread's planes of the previous tick."""
    X, Y, Z = 70, 11, 6
    cells = O.init_cells(X, Y, Z, 3, (X, Y, Z), 0, 0.35, 99)
    for threads in (1, 2, 3, 5, 6, 9):
        _check(X, Y, Z, 3, 7, 0x1C0, 0xF8, cells, 6, threads)


def test_bad_arguments():
    cells = O.init_cells(8, 8, 8, 2, (8, 8, 8), 0, 0.3, 1)
    g = O.FastGrid(cells, 8, 8, 8)
    for M, flags in ((0, 7), (16, 7), (3, 0), (3, 8)):
        with pytest.raises(RuntimeError):
            g.run(M, flags, 1, 1, 1)
    g.close()


def _sims(n, seed):
    rng = np.random.default_rng(seed)
    sims = []
    for _ in range(n):
        X, Y, Z = (int(v) for v in rng.integers(8, 36, 3))
        M = int(rng.integers(1, 8))
        flags = int(rng.choice([1, 3, 7]))
        nmax = O.lib().orc_neighbor_count(flags)
        birth = int(rng.integers(0, 1 << nmax)) & int(rng.integers(0, 1 << nmax))
        surv = int(rng.integers(0, 1 << nmax)) & int(rng.integers(0, 1 << nmax))
        sims.append(dict(X=X, Y=Y, Z=Z, M=M, flags=flags, max_ticks=int(rng.integers(5, 40)), birth=birth,
                         survival=surv, sx=max(1, X // 2), sy=max(1, Y // 2), sz=max(1, Z // 2), spawn_mode=0,
                         seed=int(rng.integers(1 << 30)), density=float(rng.uniform(0.1, 0.5))))
    return sims


def test_run_sims_matches_run_rule():
    """Whole simulations (fresh cells, manager after every tick, early
    death) against orc_run_rule: ticks run and every grid-info field."""
    sims = _sims(40, 11)
    ticks, infos, done, ups, _, hashes = O.run_sims_fast(sims, n_threads=4, hashes=True)
    assert done == len(sims)
    total = 0
    for i, sp in enumerate(sims):
        init = O.init_cells(sp["X"], sp["Y"], sp["Z"], sp["M"], (sp["sx"], sp["sy"], sp["sz"]), 0,
                            sp["density"], sp["seed"])
        t, info, _, final = O.run_rule(sp["X"], sp["Y"], sp["Z"], sp["M"], sp["flags"], sp["birth"],
                                       sp["survival"], sp["max_ticks"], init)
        assert ticks[i] == t, i
        assert int(hashes[i]) == O.canonical_hash(final, sp["X"], sp["Y"], sp["Z"]), i
        for f, _ in O.GridInfo._fields_:
            assert getattr(infos[i], f) == getattr(info, f), (i, f)
        total += t * sp["X"] * sp["Y"] * sp["Z"]
    assert ups == total


def test_run_sims_budget():
    """A spent time budget starts no simulation: those report UINT32_MAX."""
    sims = _sims(30, 12)
    ticks, _, done, _, _ = O.run_sims_fast(sims, n_threads=2, budget_s=1e-9)
    assert done == int((ticks != UINT32_MAX).sum())
    assert done < len(sims)
