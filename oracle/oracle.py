"""ctypes binding of the CPU oracle (oracle/cave_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, as the checker or the timed CPU baseline.  The
product (cave_amd) never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libcave_oracle.so")


class GridInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "totalNumberOfCellsThatDied", "totalNumberOfCellsThatWereBorn",
        "numberOfSimulationTicksSurvived", "status", "maxAliveCount",
        "finalAliveCount")]


class Snapshot(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "numberOfDeadCells", "numberOfDyingCells", "numberOfCellsBorn",
        "numberOfSurvivingCells")]


class Permutation(C.Structure):
    _fields_ = [("birthAndMaxCellStateRules", C.c_uint64),
                ("survivalAndNeighborhoodRules", C.c_uint64)]


class Sim(C.Structure):
    """ofs_sim (cpu_fast.h): one simulation of a list for the fast CPU form."""
    _fields_ = [(n, C.c_uint32) for n in ("X", "Y", "Z", "M", "flags", "max_ticks")] + \
               [("birth", C.c_uint64), ("survival", C.c_uint64)] + \
               [(n, C.c_uint32) for n in ("sx", "sy", "sz", "spawn_mode", "seed", "reserved")] + \
               [("density", C.c_double)]


class TickStats(C.Structure):
    _fields_ = [("snap", Snapshot), ("live", C.c_uint32), ("changed", C.c_uint32),
                ("bbox_min", C.c_int32 * 3), ("bbox_max", C.c_int32 * 3)]

    def as_tuple(self):
        s = self.snap
        return (s.numberOfDeadCells, s.numberOfDyingCells, s.numberOfCellsBorn,
                s.numberOfSurvivingCells, self.live, self.changed,
                tuple(self.bbox_min), tuple(self.bbox_max))


def build() -> str:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        u32, u64, p32 = C.c_uint32, C.c_uint64, C.POINTER(C.c_uint32)
        L.orc_words_per_row.restype = u32
        L.orc_cell_words.restype = C.c_size_t
        L.orc_cell_words.argtypes = [u32, u32, u32]
        L.orc_canonical_hash.restype = u64
        L.orc_canonical_hash.argtypes = [p32, u32, u32, u32]
        L.orc_wang_hash.restype = u32
        L.orc_wang_hash.argtypes = [u32]
        L.orc_cell_hash.restype = u32
        L.orc_cell_hash.argtypes = [u32, u32, u32, u32]
        L.orc_density_threshold.restype = u32
        L.orc_density_threshold.argtypes = [C.c_double]
        L.orc_init_cells.restype = None
        L.orc_init_cells.argtypes = [p32, u32, u32, u32, u32, u32, u32, u32, u32, C.c_double, u32]
        L.orc_step.restype = None
        L.orc_step.argtypes = [p32, p32, u32, u32, u32, u64, u64, C.POINTER(C.c_uint64), u32,
                               C.POINTER(Snapshot)]
        L.orc_step_stats.restype = None
        L.orc_step_stats.argtypes = [p32, p32, u32, u32, u32, u64, u64, C.POINTER(C.c_uint64), u32,
                                     C.POINTER(TickStats)]
        L.orc_manager_tick.restype = u32
        L.orc_manager_tick.argtypes = [C.POINTER(GridInfo), C.POINTER(Snapshot), C.POINTER(Permutation),
                                       u64, u64, u32, C.POINTER(GridInfo), C.POINTER(Permutation)]
        L.orc_classify.restype = u32
        L.orc_classify.argtypes = [C.POINTER(GridInfo), u32]
        L.orc_enumerate_valid_bitmasks.restype = C.c_size_t
        L.orc_enumerate_valid_bitmasks.argtypes = [u32, u32, C.POINTER(C.c_uint64), C.c_size_t]
        L.orc_rule_permutation_count.restype = u64
        L.orc_rule_permutation_count.argtypes = [u32]
        L.orc_neighbor_count.restype = u32
        L.orc_neighbor_count.argtypes = [u32]
        L.orc_search_total_permutations.restype = u64
        L.orc_search_total_permutations.argtypes = [C.c_int] * 6
        L.orc_rule_string_to_bitmask.restype = u64
        L.orc_rule_string_to_bitmask.argtypes = [C.c_char_p]
        L.orc_bitmask_to_rule_string.restype = C.c_size_t
        L.orc_bitmask_to_rule_string.argtypes = [u64, C.c_char_p, C.c_size_t]
        L.orc_run_rule.restype = u32
        L.orc_run_rule.argtypes = [u32, u32, u32, u32, u32, u64, u64, u32, p32, C.POINTER(GridInfo),
                                   C.POINTER(TickStats), p32]
        L.orc_run_search.restype = u64
        L.orc_run_search.argtypes = [u32, u32, u32, C.c_int, C.c_int, u32, u32, u32, C.c_double, u32,
                                     C.c_int, C.POINTER(C.c_uint64)]
        L.orc_erd_words.restype = C.c_size_t
        L.orc_erd_words.argtypes = [u32, u32, u32]
        L.orc_erd_neighbor_count.restype = u32
        L.orc_erd_neighbor_count.argtypes = [u32]
        L.orc_step_erd.restype = None
        L.orc_step_erd.argtypes = [p32, p32, u32, u32, u32, u64, u64, C.POINTER(C.c_uint64), u32,
                                   C.POINTER(TickStats)]
        L.orc_init_cells_erd.restype = None
        L.orc_init_cells_erd.argtypes = [p32, u32, u32, u32, u32, u32, u32, u32, u32, C.c_double, u32]
        L.orc_erd_canonical_hash.restype = u64
        L.orc_erd_canonical_hash.argtypes = [p32, u32, u32, u32]
        L.orc_run_rule_erd.restype = u32
        L.orc_run_rule_erd.argtypes = [u32, u32, u32, u32, u32, u64, u64, u32, p32, C.POINTER(GridInfo),
                                       C.POINTER(TickStats), p32]
        L.orc_run_search_shape.restype = u64
        L.orc_run_search_shape.argtypes = [u32, u32, u32, u32, C.c_int, C.c_int, u32, u32, u32, C.c_double, u32,
                                           C.c_int, C.POINTER(C.c_uint64), C.c_int]
        L.orc_step_render.restype = None
        L.orc_step_render.argtypes = [p32, p32, u32, u32, u32, u64, u64, C.POINTER(TickStats)]
        L.orc_render_density_threshold.restype = u32
        L.orc_render_density_threshold.argtypes = [C.c_double, C.POINTER(C.c_int)]
        L.orc_init_cells_render.restype = None
        L.orc_init_cells_render.argtypes = [p32, u32, u32, u32, u32, u32, u32, u32, u32, C.c_double]
        vp = C.c_void_p
        L.ofs_create.restype = vp
        L.ofs_create.argtypes = [u32, u32, u32, p32]
        L.ofs_destroy.restype = None
        L.ofs_destroy.argtypes = [vp]
        L.ofs_read.restype = None
        L.ofs_read.argtypes = [vp, p32]
        L.ofs_run.restype = C.c_int
        L.ofs_run.argtypes = [vp, u32, u32, u64, u64, u32, C.c_int, C.POINTER(TickStats)]
        L.ofs_run_sims.restype = C.c_size_t
        L.ofs_run_sims.argtypes = [C.POINTER(Sim), C.c_size_t, C.c_int, C.c_double, p32, C.POINTER(GridInfo),
                                   C.POINTER(C.c_uint64)]
        L.ofs_run_sims_hash.restype = C.c_size_t
        L.ofs_run_sims_hash.argtypes = [C.POINTER(Sim), C.c_size_t, C.c_int, C.c_double, p32, C.POINTER(GridInfo),
                                        C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        _lib = L
    return _lib


def _p32(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def cell_words(X, Y, Z) -> int:
    return int(lib().orc_cell_words(X, Y, Z))


def init_cells(X, Y, Z, M, spawn, spawn_mode, density, seed) -> np.ndarray:
    a = np.zeros(cell_words(X, Y, Z), dtype=np.uint32)
    sx, sy, sz = spawn
    lib().orc_init_cells(_p32(a), X, Y, Z, M, sx, sy, sz, spawn_mode, float(density), seed)
    return a


# ---- ERD shape (linear layout) ----
def erd_words(X, Y, Z) -> int:
    return int(lib().orc_erd_words(X, Y, Z))


def step_erd(cells, X, Y, Z, word0, word1, valid_bitmasks=None, max_rule_bits=0):
    out = np.zeros_like(cells)
    st = TickStats()
    vb = None
    if valid_bitmasks is not None:
        vb_arr = np.ascontiguousarray(valid_bitmasks, dtype=np.uint64)
        vb = vb_arr.ctypes.data_as(C.POINTER(C.c_uint64))
    lib().orc_step_erd(_p32(cells), _p32(out), X, Y, Z, word0, word1, vb, max_rule_bits, C.byref(st))
    return out, st


def init_cells_erd(X, Y, Z, M, spawn, spawn_mode, density, seed) -> np.ndarray:
    a = np.zeros(erd_words(X, Y, Z), dtype=np.uint32)
    sx, sy, sz = spawn
    lib().orc_init_cells_erd(_p32(a), X, Y, Z, M, sx, sy, sz, spawn_mode, float(density), seed)
    return a


def erd_hash(cells, X, Y, Z) -> int:
    return int(lib().orc_erd_canonical_hash(_p32(cells), X, Y, Z))


def run_rule_erd(X, Y, Z, M, flags, birth, survival, max_ticks, initial):
    info = GridInfo()
    per_tick = (TickStats * max_ticks)()
    final = np.zeros_like(initial)
    t = lib().orc_run_rule_erd(X, Y, Z, M, flags, birth, survival, max_ticks, _p32(initial), C.byref(info),
                               per_tick, _p32(final))
    return int(t), info, [per_tick[i] for i in range(min(t, max_ticks))], final


def run_search_shape(grid, shape, sel_flags, max_rule_bits, min_m, max_m, max_ticks, spawn, spawn_mode,
                     density, seed, n_threads=8, legacy_death=0):
    counts = (C.c_uint64 * 3)()
    n = lib().orc_run_search_shape(grid, shape, sel_flags, max_rule_bits, min_m, max_m, max_ticks, spawn,
                                   spawn_mode, float(density), seed, n_threads, counts, legacy_death)
    return int(n), (int(counts[0]), int(counts[1]), int(counts[2]))


def step_render(cells, X, Y, Z, word0, word1):
    """The headless render-path step (wraps every axis when word1 has the W flag)."""
    out = np.zeros_like(cells)
    st = TickStats()
    lib().orc_step_render(_p32(cells), _p32(out), X, Y, Z, word0, word1, C.byref(st))
    return out, st


def render_density_threshold(density):
    """(threshold, all_alive) of the render path's random spawn."""
    a = C.c_int(0)
    t = lib().orc_render_density_threshold(float(density), C.byref(a))
    return int(t), bool(a.value)


def init_cells_render(X, Y, Z, M, spawn, spawn_mode, density) -> np.ndarray:
    a = np.zeros(cell_words(X, Y, Z), dtype=np.uint32)
    sx, sy, sz = spawn
    lib().orc_init_cells_render(_p32(a), X, Y, Z, M, sx, sy, sz, spawn_mode, float(density))
    return a


def step(cells, X, Y, Z, word0, word1, valid_bitmasks=None, max_rule_bits=0):
    out = np.zeros_like(cells)
    st = TickStats()
    vb = None
    if valid_bitmasks is not None:
        vb_arr = np.ascontiguousarray(valid_bitmasks, dtype=np.uint64)
        vb = vb_arr.ctypes.data_as(C.POINTER(C.c_uint64))
    lib().orc_step_stats(_p32(cells), _p32(out), X, Y, Z, word0, word1, vb, max_rule_bits,
                         C.byref(st))
    return out, st


def canonical_hash(cells, X, Y, Z) -> int:
    return int(lib().orc_canonical_hash(_p32(cells), X, Y, Z))


def enumerate_valid_bitmasks(N, K):
    n = lib().orc_enumerate_valid_bitmasks(N, K, None, 0)
    out = (C.c_uint64 * max(n, 1))()
    lib().orc_enumerate_valid_bitmasks(N, K, out, n)
    return [int(out[i]) for i in range(n)]


def rule_string_to_bitmask(s: str) -> int:
    return int(lib().orc_rule_string_to_bitmask(s.encode()))


def bitmask_to_rule_string(m: int) -> str:
    buf = C.create_string_buffer(512)
    lib().orc_bitmask_to_rule_string(m, buf, 512)
    return buf.value.decode()


def run_rule(X, Y, Z, M, flags, birth, survival, max_ticks, initial):
    info = GridInfo()
    per_tick = (TickStats * max_ticks)()
    final = np.zeros_like(initial)
    t = lib().orc_run_rule(X, Y, Z, M, flags, birth, survival, max_ticks, _p32(initial),
                           C.byref(info), per_tick, _p32(final))
    return int(t), info, [per_tick[i] for i in range(min(t, max_ticks))], final


def run_search(grid, sel_flags, max_rule_bits, min_m, max_m, max_ticks, spawn, spawn_mode,
               density, seed, n_threads=8):
    counts = (C.c_uint64 * 3)()
    n = lib().orc_run_search(grid, sel_flags, max_rule_bits, min_m, max_m, max_ticks, spawn,
                             spawn_mode, float(density), seed, n_threads, counts)
    return int(n), (int(counts[0]), int(counts[1]), int(counts[2]))


# ---- the fast CPU form (cpu_fast.c): bit-sliced, threaded ----
class FastGrid:
    """A grid on the fast CPU form.  Same step as step(); the cells go in and
    come out in the packed layout."""

    def __init__(self, cells, X, Y, Z):
        self.X, self.Y, self.Z = X, Y, Z
        self._h = lib().ofs_create(X, Y, Z, _p32(np.ascontiguousarray(cells, dtype=np.uint32)))
        if not self._h:
            raise MemoryError("ofs_create failed")

    def run(self, M, flags, birth, survival, ticks, n_threads=1, stats=True):
        per = (TickStats * ticks)() if stats else None
        rc = lib().ofs_run(self._h, M, flags, birth, survival, ticks, n_threads, per)
        if rc != 0:
            raise RuntimeError(f"ofs_run failed ({rc})")
        return [per[i] for i in range(ticks)] if stats else None

    def cells(self):
        out = np.zeros(cell_words(self.X, self.Y, self.Z), dtype=np.uint32)
        lib().ofs_read(self._h, _p32(out))
        return out

    def close(self):
        if self._h:
            lib().ofs_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()


SIM_NP = np.dtype([(n, np.uint32) for n in ("X", "Y", "Z", "M", "flags", "max_ticks")] +
                  [("birth", np.uint64), ("survival", np.uint64)] +
                  [(n, np.uint32) for n in ("sx", "sy", "sz", "spawn_mode", "seed", "reserved")] +
                  [("density", np.float64)])
assert SIM_NP.itemsize == C.sizeof(Sim)


def run_sims_fast(sims, n_threads=1, budget_s=0.0, hashes=False):
    """Runs a list of simulations on the fast CPU form: `sims` is a SIM_NP
    array or a list of dicts with its fields.  Returns (ticks per sim,
    UINT32_MAX where not run; grid infos; number run; cell-updates executed;
    seconds inside the call), and with hashes=True a sixth item: the final
    cells' canonical hash per simulation."""
    if not isinstance(sims, np.ndarray):
        arr = np.zeros(len(sims), dtype=SIM_NP)
        for i, sp in enumerate(sims):
            for k, v in sp.items():
                arr[k][i] = v
        sims = arr
    sims = np.ascontiguousarray(sims, dtype=SIM_NP)
    n = len(sims)
    ticks = np.zeros(max(n, 1), dtype=np.uint32)
    infos = (GridInfo * max(n, 1))()
    ups = C.c_uint64(0)
    import time

    hs = np.zeros(max(n, 1), dtype=np.uint64) if hashes else None
    t0 = time.perf_counter()
    done = lib().ofs_run_sims_hash(sims.ctypes.data_as(C.POINTER(Sim)), n, n_threads, float(budget_s), _p32(ticks),
                                   infos, C.byref(ups),
                                   hs.ctypes.data_as(C.POINTER(C.c_uint64)) if hashes else None)
    dt = time.perf_counter() - t0
    out = (ticks[:n], [infos[i] for i in range(n)], int(done), int(ups.value), dt)
    return out + (hs[:n],) if hashes else out
