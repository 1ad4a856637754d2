"""cave_amd -- MI355X-native CAVE step and rule-search engine (Python mirror).

The compute path is the native library cave_amd/_lib/libcave_amd.so (HIP
kernels for gfx950 behind the C-ABI declared in include/cave.h).  This module
is a thin ctypes mirror of that ABI, shaped like the reference's own
interfaces (SearchSystem, the headless step loop).  There is no Python or CPU
fallback: if the library is missing, loading it raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "libcave_amd.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "cave.h")

CAVE_OK, CAVE_E_INVALID, CAVE_E_UNSUPPORTED, CAVE_E_NOMEM, CAVE_E_DEVICE, CAVE_E_BUDGET, CAVE_E_STATE = range(7)
FACES, EDGES, CORNERS, WRAP = 1, 2, 4, 8
SPAWN_RANDOM, SPAWN_FILLED, SPAWN_BUFFER, SPAWN_RENDER_RANDOM = 0, 1, 2, 3
KERNEL_AUTO, KERNEL_GATHER, KERNEL_STREAM, KERNEL_STREAM_NIBBLE, KERNEL_ERD = 0, 1, 2, 3, 4
ENGINE_AUTO, ENGINE_PERSISTENT, ENGINE_PLAIN = 0, 1, 2
SIM_NO_EARLY_EXIT, SIM_FINAL_HASH, SIM_LEGACY_DEATH, SIM_SHAPE_ERD, SIM_FULL_GRID = 1, 2, 4, 8, 16
SHAPE_CUBE, SHAPE_ERD = 0, 1


class CaveError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cave error {code}: {msg}")
        self.code = code


# ---- structs (byte-identical to include/cave.h) ----
class SimulationParameters(C.Structure):
    _fields_ = [("birthAndMaxCellStateRules", C.c_uint64),
                ("survivalAndNeighborhoodRules", C.c_uint64),
                ("shapeAndGridConfiguration", C.c_uint64)]


class GridSnapshot(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("numberOfDeadCells", "numberOfDyingCells",
                                          "numberOfCellsBorn", "numberOfSurvivingCells")]


class GridInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "totalNumberOfCellsThatDied", "totalNumberOfCellsThatWereBorn",
        "numberOfSimulationTicksSurvived", "status", "maxAliveCount", "finalAliveCount")]


class ChunkConfiguration(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in (
        "shapeId", "gridDimensionX", "gridDimensionY", "gridDimensionZ",
        "spawnAreaDimensionX", "spawnAreaDimensionY", "spawnAreaDimensionZ",
        "spawnPlacementId", "spawnRegionShapeId", "spawnSizeIndex", "spawnDensityIndex",
        "runSeedIndex", "partitionIndex", "partitionCount")]


class TickStats(C.Structure):
    _fields_ = [("snap", GridSnapshot), ("live", C.c_uint32), ("changed", C.c_uint32),
                ("bbox_min", C.c_int32 * 3), ("bbox_max", C.c_int32 * 3)]

    def as_tuple(self):
        s = self.snap
        return (s.numberOfDeadCells, s.numberOfDyingCells, s.numberOfCellsBorn,
                s.numberOfSurvivingCells, self.live, self.changed,
                tuple(self.bbox_min), tuple(self.bbox_max))


class StepSpec(C.Structure):
    _fields_ = [("dims", C.c_uint32 * 3), ("params", SimulationParameters),
                ("spawn", C.c_uint32 * 3), ("spawn_mode", C.c_uint32), ("density", C.c_double),
                ("seed", C.c_uint32), ("initial_cells", C.POINTER(C.c_uint32)),
                ("kernel", C.c_uint32), ("reserved", C.c_uint32)]


class SimSpec(C.Structure):
    _fields_ = [("dims", C.c_uint32 * 3), ("max_cell_state", C.c_uint32),
                ("neighborhood_flags", C.c_uint32), ("ticks", C.c_uint32),
                ("birth", C.c_uint64), ("survival", C.c_uint64),
                ("spawn", C.c_uint32 * 3), ("spawn_mode", C.c_uint32), ("density", C.c_double),
                ("seed", C.c_uint32), ("flags", C.c_uint32)]


class SimResult(C.Structure):
    _fields_ = [("info", GridInfo), ("classification", C.c_uint32), ("ticks_run", C.c_uint32),
                ("last", TickStats), ("final_hash", C.c_uint64)]


class BatchOptions(C.Structure):
    _fields_ = [("engine", C.c_uint32), ("block_threads", C.c_uint32),
                ("per_tick", C.POINTER(TickStats)), ("per_tick_stride", C.c_uint32),
                ("max_parts", C.c_uint32), ("device_ms", C.c_double),
                ("timing", C.c_void_p)]


class SearchSpec(C.Structure):
    _fields_ = [("chunk", ChunkConfiguration), ("neighborhood_flags", C.c_uint32),
                ("max_rule_bits", C.c_uint32), ("min_max_cell_state", C.c_int32),
                ("max_max_cell_state", C.c_int32), ("max_ticks", C.c_uint32),
                ("sim_flags", C.c_uint32), ("output_folder", C.c_char_p), ("file_prefix", C.c_char_p)]


class SearchResult(C.Structure):
    _fields_ = [("permutation", SimulationParameters), ("ticksSurvived", C.c_uint32),
                ("totalBirths", C.c_uint32), ("totalDeaths", C.c_uint32),
                ("maxAliveCount", C.c_uint32), ("finalAliveCount", C.c_uint32),
                ("category", C.c_uint32)]


class SearchProgress(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("total", "completed", "viable", "possible", "unviable")]


class ServeOptions(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_uint64), ("max_inflight", C.c_uint32), ("split_target_us", C.c_uint32),
                ("idle_exit_us", C.c_uint32), ("max_parts", C.c_uint32)]


class ServeCompletion(C.Structure):
    _fields_ = [("request_id", C.c_uint64), ("result", SimResult), ("submit_s", C.c_double),
                ("admit_s", C.c_double), ("finish_s", C.c_double), ("run_ms", C.c_double),
                ("blocks", C.c_uint32), ("status", C.c_int32)]


_lib = None

# symbol -> (restype, argtypes)
_u32, _u64, _i32, _sz, _vp = C.c_uint32, C.c_uint64, C.c_int32, C.c_size_t, C.c_void_p
_SIGS = {
    "cave_abi_version": (C.c_int, []),
    "cave_last_error": (C.c_char_p, []),
    "cave_device_info": (C.c_int, [C.c_char_p, _sz, C.POINTER(C.c_int), C.POINTER(_u64)]),
    "cave_rule_string_to_bitmask": (_u64, [C.c_char_p]),
    "cave_bitmask_to_rule_string": (_sz, [_u64, C.c_char_p, _sz]),
    "cave_enumerate_valid_bitmasks": (_sz, [_u32, _u32, C.POINTER(_u64), _sz]),
    "cave_cell_words": (_sz, [_u32, _u32, _u32]),
    "cave_splitmix64": (_u64, [C.POINTER(_u64)]),
    "cave_random_rule": (_u64, [C.POINTER(_u64), _u32, C.c_double]),
    "cave_step_create": (C.c_int, [C.POINTER(StepSpec), C.POINTER(_vp)]),
    "cave_step_run": (C.c_int, [_vp, _u32, C.POINTER(TickStats), C.POINTER(C.c_double)]),
    "cave_step_download": (C.c_int, [_vp, C.POINTER(_u32), _sz]),
    "cave_step_upload": (C.c_int, [_vp, C.POINTER(_u32), _sz]),
    "cave_step_set_params": (C.c_int, [_vp, C.POINTER(SimulationParameters)]),
    "cave_step_kernel": (C.c_int, [_vp]),
    "cave_step_destroy": (C.c_int, [_vp]),
    "cave_step_run_static": (C.c_int, [C.POINTER(SimulationParameters), C.POINTER(_u32), _u32, _vp, _vp, _u32,
                                       _vp, _vp]),
    "cave_batch_run": (C.c_int, [C.POINTER(SimSpec), _sz, C.POINTER(SimResult), C.POINTER(BatchOptions)]),
    "cave_batch_tuning_report": (C.c_int, [C.c_char_p, _sz]),
    "cave_render_bytes": (_sz, [_u32, _u32, _u32, _u32]),
    "cave_erd_cell_words": (_sz, [_u32, _u32, _u32]),
    "cave_render_to_packed": (C.c_int, [_u32, _vp, _u32, _u32, _u32, C.POINTER(_u32)]),
    "cave_packed_to_render": (C.c_int, [_u32, C.POINTER(_u32), _u32, _u32, _u32, _vp]),
    "cave_search_create": (C.c_int, [C.POINTER(SearchSpec), C.POINTER(_vp)]),
    "cave_search_step": (C.c_int, [_vp, _u64]),
    "cave_search_step_group": (C.c_int, [C.POINTER(_vp), _sz, _u64]),
    "cave_estimate_chunk_vram_bytes": (_u64, [_u32, _u32, _u32, _u32]),
    "cave_aggregate_create": (C.c_int, [C.POINTER(_vp)]),
    "cave_aggregate_record": (C.c_int, [_vp, C.POINTER(SimulationParameters), _u32, _u32, _u32]),
    "cave_aggregate_record_batch": (C.c_int, [_vp, C.POINTER(SimSpec), C.POINTER(SimResult), _sz]),
    "cave_aggregate_rule_count": (_sz, [_vp]),
    "cave_aggregate_write_jsonl": (C.c_int, [_vp, C.c_char_p]),
    "cave_aggregate_destroy": (C.c_int, [_vp]),
    "cave_search_is_running": (C.c_int, [_vp]),
    "cave_search_progress": (C.c_int, [_vp, C.POINTER(SearchProgress)]),
    "cave_search_drain": (C.c_int, [_vp, C.POINTER(SearchResult), _sz, C.POINTER(_sz)]),
    "cave_search_write_summary": (C.c_int, [_vp, C.c_char_p, _sz]),
    "cave_search_destroy": (C.c_int, [_vp]),
    "cave_serve_create": (C.c_int, [C.POINTER(ServeOptions), C.POINTER(_vp)]),
    "cave_serve_submit": (C.c_int, [_vp, C.POINTER(SimSpec), _u64]),
    "cave_serve_poll": (C.c_int, [_vp, C.POINTER(ServeCompletion), _sz, C.POINTER(_sz)]),
    "cave_serve_drain": (C.c_int, [_vp]),
    "cave_serve_stats": (C.c_int, [_vp, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
    "cave_serve_destroy": (C.c_int, [_vp]),
}


def lib():
    """Load libcave_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m cave_amd.build` "
                              "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != CAVE_OK:
        raise CaveError(rc, lib().cave_last_error().decode(errors="replace"))


def _p32(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


# ---- host helpers ----
def rule_string_to_bitmask(s: str) -> int:
    return int(lib().cave_rule_string_to_bitmask(s.encode()))


def bitmask_to_rule_string(m: int) -> str:
    buf = C.create_string_buffer(512)
    lib().cave_bitmask_to_rule_string(m, buf, 512)
    return buf.value.decode()


def enumerate_valid_bitmasks(n: int, k: int) -> list[int]:
    cnt = lib().cave_enumerate_valid_bitmasks(n, k, None, 0)
    out = (C.c_uint64 * max(cnt, 1))()
    lib().cave_enumerate_valid_bitmasks(n, k, out, cnt)
    return [int(out[i]) for i in range(cnt)]


def erd_cell_words(X: int, Y: int, Z: int) -> int:
    return int(lib().cave_erd_cell_words(X, Y, Z))


def cell_words(X: int, Y: int, Z: int) -> int:
    return int(lib().cave_cell_words(X, Y, Z))


class Rng:
    """splitmix64 stream shared with the native library (cave_splitmix64)."""

    def __init__(self, seed: int):
        self.state = C.c_uint64(seed)

    def next(self) -> int:
        return int(lib().cave_splitmix64(C.byref(self.state)))

    def rule(self, n_counts: int, p: float) -> int:
        return int(lib().cave_random_rule(C.byref(self.state), n_counts, p))

    def uniform(self) -> float:
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)


def device_info():
    name = C.create_string_buffer(256)
    cu = C.c_int(0)
    mem = C.c_uint64(0)
    check(lib().cave_device_info(name, 256, C.byref(cu), C.byref(mem)))
    return name.value.decode(), cu.value, mem.value


def encode_params(M: int, flags: int, birth: int, survival: int, shape: int = 0) -> SimulationParameters:
    """SimulationParameters in raw-bitmask form (src/common/structs.h:89-99,
    src/vulkan/systems/guiSystem.cpp:370-388)."""
    p = SimulationParameters()
    p.birthAndMaxCellStateRules = (((M - 1) & 0xF) << 60) | (birth & ((1 << 60) - 1))
    p.survivalAndNeighborhoodRules = ((flags & 0xF) << 60) | (survival & ((1 << 60) - 1))
    p.shapeAndGridConfiguration = shape & 0x1F
    return p


# ---- single-grid step ----
class Step:
    """One resident grid advanced tick by tick on the GPU (headless step path)."""

    def __init__(self, dims, M, flags, birth, survival, spawn=None, spawn_mode=SPAWN_RANDOM,
                 density=0.5, seed=0, initial_cells=None, kernel=KERNEL_AUTO, shape=SHAPE_CUBE):
        spec = StepSpec()
        spec.dims[:] = list(dims)
        spec.params = encode_params(M, flags, birth, survival, shape)
        self.shape = shape
        spec.spawn[:] = list(spawn if spawn is not None else dims)
        spec.density = float(density)
        spec.seed = seed
        spec.kernel = kernel
        self._keep = None
        if initial_cells is not None:
            arr = np.ascontiguousarray(initial_cells, dtype=np.uint32)
            self._keep = arr
            spec.spawn_mode = SPAWN_BUFFER
            spec.initial_cells = _p32(arr)
        else:
            spec.spawn_mode = spawn_mode
        self.dims = tuple(dims)
        h = C.c_void_p()
        check(lib().cave_step_create(C.byref(spec), C.byref(h)))
        self._h = h
        self._keep = None

    @property
    def kernel(self) -> int:
        return int(lib().cave_step_kernel(self._h))

    def run(self, ticks: int, stats: bool = True):
        """Advances `ticks` ticks; returns (per-tick TickStats list, device ms)."""
        arr = (TickStats * max(ticks, 1))() if stats else None
        ms = C.c_double(0)
        check(lib().cave_step_run(self._h, ticks, arr, C.byref(ms)))
        return ([arr[i] for i in range(ticks)] if stats else None), ms.value

    def cells(self) -> np.ndarray:
        """Current cells: the packed search layout, or the linear layout for ERD."""
        X, Y, Z = self.dims
        n = erd_cell_words(X, Y, Z) if self.shape == SHAPE_ERD else cell_words(X, Y, Z)
        out = np.zeros(n, dtype=np.uint32)
        check(lib().cave_step_download(self._h, _p32(out), out.size))
        return out

    def upload(self, cells: np.ndarray) -> None:
        arr = np.ascontiguousarray(cells, dtype=np.uint32)
        check(lib().cave_step_upload(self._h, _p32(arr), arr.size))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cave_step_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def step_run_static(dims, M, flags, birth, survival, cells_a: int, cells_b: int, ticks: int,
                    per_tick: int = 0, stream: int = 0, kernel=KERNEL_AUTO) -> None:
    """cave_step_run_static on device pointers (integers): `ticks` ticks on
    caller-owned buffers, no device allocation; the state ends in cells_a
    when ticks is even, in cells_b when odd.  per_tick: device pointer to
    `ticks` TickStats (0 = none).  Queued on `stream`, not waited for."""
    p = encode_params(M, flags, birth, survival)
    d = (C.c_uint32 * 3)(*dims)
    check(lib().cave_step_run_static(C.byref(p), d, kernel, cells_a, cells_b, ticks, per_tick or None,
                                     stream or None))


# ---- batch ----
@dataclass
class Sim:
    dims: tuple
    M: int
    flags: int
    ticks: int
    birth: int
    survival: int
    spawn: tuple
    spawn_mode: int = SPAWN_RANDOM
    density: float = 0.5
    seed: int = 0
    sim_flags: int = 0

    def to_c(self, s: SimSpec) -> None:
        s.dims[:] = list(self.dims)
        s.max_cell_state = self.M
        s.neighborhood_flags = self.flags
        s.ticks = self.ticks
        s.birth = self.birth
        s.survival = self.survival
        s.spawn[:] = list(self.spawn)
        s.spawn_mode = self.spawn_mode
        s.density = float(self.density)
        s.seed = self.seed
        s.flags = self.sim_flags


# numpy mirrors of CaveSimSpec / CaveSimResult for large batches
SIM_DTYPE = np.dtype({
    "names": ["dims", "M", "flags", "ticks", "birth", "survival", "spawn", "spawn_mode", "density",
              "seed", "sim_flags"],
    "formats": [("<u4", 3), "<u4", "<u4", "<u4", "<u8", "<u8", ("<u4", 3), "<u4", "<f8", "<u4", "<u4"],
    "offsets": [0, 12, 16, 20, 24, 32, 40, 52, 56, 64, 68],
    "itemsize": 72})
RESULT_DTYPE = np.dtype({
    "names": ["died", "born", "ticks_survived", "status", "max_alive", "final_alive", "classification",
              "ticks_run", "last", "final_hash"],
    "formats": ["<u4", "<u4", "<u4", "<u4", "<u4", "<u4", "<u4", "<u4", ("<u4", 12), "<u8"],
    "offsets": [0, 4, 8, 12, 16, 20, 24, 28, 32, 80],
    "itemsize": 88})


TIMING_DTYPE = np.dtype([("start_us", "<f8"), ("end_us", "<f8"), ("blocks", "<u4"), ("items_per_thread", "<u4"),
                         ("computed_cells", "<u8")])


def batch_run_np(specs: np.ndarray, engine=ENGINE_AUTO, block_threads=0, max_parts=0, timing=False):
    """Evaluates a SIM_DTYPE array; returns (RESULT_DTYPE array, device ms),
    plus a TIMING_DTYPE array (device start / end microseconds, blocks, items
    per thread, cells computed per simulation) and the block size used when timing is set
    (persistent engine)."""
    specs = np.ascontiguousarray(specs, dtype=SIM_DTYPE)
    res = np.zeros(len(specs), dtype=RESULT_DTYPE)
    opt = BatchOptions()
    opt.engine = engine
    opt.block_threads = block_threads
    opt.max_parts = max_parts
    tm = None
    if timing:
        tm = np.zeros(len(specs), dtype=TIMING_DTYPE)
        opt.timing = tm.ctypes.data
    check(lib().cave_batch_run(specs.ctypes.data_as(C.POINTER(SimSpec)), len(specs),
                               res.ctypes.data_as(C.POINTER(SimResult)), C.byref(opt)))
    if timing:
        return res, opt.device_ms, tm, opt.block_threads
    return res, opt.device_ms


RENDER_IMAGE, RENDER_TILED = 0, 1


def render_to_packed(layout: int, render: np.ndarray, X: int, Y: int, Z: int) -> np.ndarray:
    """Render-path storage (image bytes or tiled u32 buffer) -> packed search layout."""
    src = np.ascontiguousarray(render)
    out = np.zeros(cell_words(X, Y, Z), dtype=np.uint32)
    check(lib().cave_render_to_packed(layout, src.ctypes.data, X, Y, Z, _p32(out)))
    return out


def packed_to_render(layout: int, packed: np.ndarray, X: int, Y: int, Z: int) -> np.ndarray:
    """Packed search layout -> render-path storage (uint8 image or uint32 tiled buffer)."""
    n = int(lib().cave_render_bytes(layout, X, Y, Z))
    out = np.zeros(n if layout == RENDER_IMAGE else n // 4, dtype=np.uint8 if layout == RENDER_IMAGE else np.uint32)
    src = np.ascontiguousarray(packed, dtype=np.uint32)
    check(lib().cave_packed_to_render(layout, _p32(src), X, Y, Z, out.ctypes.data))
    return out


def tuning_report():
    """The persistent engine's autotuner table (list of dicts) measured so far
    in this process."""
    import json

    buf = C.create_string_buffer(1 << 20)
    check(lib().cave_batch_tuning_report(buf, len(buf)))
    return json.loads(buf.value.decode())


def batch_run(sims, engine=ENGINE_AUTO, per_tick=False, block_threads=0, max_parts=0):
    """Evaluates every Sim; returns (results, per-tick stats rows or None, device ms)."""
    n = len(sims)
    specs = (SimSpec * max(n, 1))()
    for i, s in enumerate(sims):
        s.to_c(specs[i])
    res = (SimResult * max(n, 1))()
    opt = BatchOptions()
    opt.engine = engine
    opt.block_threads = block_threads
    opt.max_parts = max_parts
    stride = max((s.ticks for s in sims), default=1)
    pt = None
    if per_tick:
        pt = (TickStats * max(n * stride, 1))()
        opt.per_tick = C.cast(pt, C.POINTER(TickStats))
        opt.per_tick_stride = stride
    check(lib().cave_batch_run(specs, n, res, C.byref(opt)))
    rows = None
    if per_tick:
        rows = [[pt[i * stride + t] for t in range(res[i].ticks_run)] for i in range(n)]
    return [res[i] for i in range(n)], rows, opt.device_ms


# ---- reference-compatible search chunk (SearchSystem) ----
class Search:
    """One --mode search chunk: every (maxCS, neighbourhood subset) slot and
    every (birth, survival) permutation of its partition, evaluated in batches
    on the GPU, streamed to <folder><prefix>_{viable,possible,unviable}.jsonl."""

    def __init__(self, grid, flags, max_rule_bits, min_cs, max_cs, ticks, spawn=11, spawn_mode=SPAWN_FILLED,
                 density_index=7, seed_index=0, partition_index=0, partition_count=1, output_folder=None,
                 file_prefix="", sim_flags=0, shape=SHAPE_CUBE):
        g = grid if isinstance(grid, (tuple, list)) else (grid, grid, grid)
        sp = spawn if isinstance(spawn, (tuple, list)) else (spawn, spawn, spawn)
        s = SearchSpec()
        c = s.chunk
        c.shapeId = shape
        c.gridDimensionX, c.gridDimensionY, c.gridDimensionZ = g
        c.spawnAreaDimensionX, c.spawnAreaDimensionY, c.spawnAreaDimensionZ = [min(a, b) for a, b in zip(sp, g)]
        c.spawnRegionShapeId = spawn_mode
        c.spawnDensityIndex = density_index
        c.runSeedIndex = seed_index
        c.partitionIndex = partition_index
        c.partitionCount = partition_count
        s.neighborhood_flags = flags
        s.max_rule_bits = max_rule_bits
        s.min_max_cell_state = min_cs
        s.max_max_cell_state = max_cs
        s.max_ticks = ticks
        s.sim_flags = sim_flags
        self._folder = output_folder.encode() if output_folder else None
        self._prefix = file_prefix.encode()
        s.output_folder = self._folder
        s.file_prefix = self._prefix
        h = C.c_void_p()
        check(lib().cave_search_create(C.byref(s), C.byref(h)))
        self._h = h

    def step(self, max_evals: int = 0) -> None:
        check(lib().cave_search_step(self._h, max_evals))

    def run(self) -> None:
        self.step(0)

    @property
    def running(self) -> bool:
        return bool(lib().cave_search_is_running(self._h))

    def progress(self) -> dict:
        p = SearchProgress()
        check(lib().cave_search_progress(self._h, C.byref(p)))
        return {n: getattr(p, n) for n, _ in SearchProgress._fields_}

    def drain(self, cap: int = 1 << 16):
        out = []
        n = C.c_size_t(0)
        while True:
            buf = (SearchResult * cap)()  # fresh buffer: the returned structs view its memory
            check(lib().cave_search_drain(self._h, buf, cap, C.byref(n)))
            out.extend(buf[i] for i in range(n.value))
            if n.value < cap:
                return out

    def summary(self) -> str:
        buf = C.create_string_buffer(4096)
        check(lib().cave_search_write_summary(self._h, buf, 4096))
        return buf.value.decode()

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cave_search_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Aggregate:
    """RuleAggregator (src/common/ruleAggregator.h:24-135): results grouped by
    rule, written as one JSONL record per rule."""

    def __init__(self):
        h = C.c_void_p()
        check(lib().cave_aggregate_create(C.byref(h)))
        self._h = h

    def record(self, params: SimulationParameters, ticks_survived: int, category: int, shape_id: int = 0) -> None:
        check(lib().cave_aggregate_record(self._h, C.byref(params), ticks_survived, category, shape_id))

    def record_batch(self, specs: np.ndarray, results: np.ndarray) -> None:
        """SIM_DTYPE specs and their RESULT_DTYPE results."""
        specs = np.ascontiguousarray(specs, dtype=SIM_DTYPE)
        results = np.ascontiguousarray(results, dtype=RESULT_DTYPE)
        assert len(specs) == len(results)
        check(lib().cave_aggregate_record_batch(self._h, specs.ctypes.data_as(C.POINTER(SimSpec)),
                                                results.ctypes.data_as(C.POINTER(SimResult)), len(specs)))

    def __len__(self) -> int:
        return int(lib().cave_aggregate_rule_count(self._h))

    def write_jsonl(self, path: str) -> None:
        check(lib().cave_aggregate_write_jsonl(self._h, path.encode()))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cave_aggregate_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def estimate_chunk_vram_bytes(X: int, Y: int, Z: int, slot_count: int) -> int:
    """PermutationMath::EstimateChunkVramBytes (permutationMath.h:144-153)."""
    return int(lib().cave_estimate_chunk_vram_bytes(X, Y, Z, slot_count))


def search_step_group(chunks, max_evals: int = 0) -> None:
    """Evaluates the pending rules of several Search chunks in shared launches."""
    arr = (C.c_void_p * max(len(chunks), 1))(*[c._h for c in chunks])
    check(lib().cave_search_step_group(arr, len(chunks), max_evals))


# ---- continuous-batching request server ----
class Server:
    """Resident persistent kernel fed with rule-evaluation requests as they
    arrive (cave_serve_*).  submit() queues and admits; poll() returns the
    finished ones as ServeCompletion records (copies)."""

    def __init__(self, hbm_budget_bytes=0, max_inflight=0, split_target_us=0, idle_exit_us=0, max_parts=0):
        o = ServeOptions()
        o.max_parts = max_parts
        o.hbm_budget_bytes = hbm_budget_bytes
        o.max_inflight = max_inflight
        o.split_target_us = split_target_us
        o.idle_exit_us = idle_exit_us
        h = C.c_void_p()
        check(lib().cave_serve_create(C.byref(o), C.byref(h)))
        self._h = h
        self._buf = (ServeCompletion * 4096)()
        self._spec = SimSpec()

    def submit(self, sim, request_id: int) -> None:
        """sim: a Sim or one SIM_DTYPE record."""
        if isinstance(sim, Sim):
            sim.to_c(self._spec)
            spec = self._spec
        else:
            rec = np.ascontiguousarray(np.asarray(sim, dtype=SIM_DTYPE).reshape(1))
            spec = SimSpec.from_buffer_copy(rec.tobytes())
        check(lib().cave_serve_submit(self._h, C.byref(spec), request_id))

    def submit_row(self, specs: np.ndarray, i: int, request_id: int) -> None:
        """Submits row i of a contiguous SIM_DTYPE array (no copy)."""
        ptr = C.cast(specs.ctypes.data + i * SIM_DTYPE.itemsize, C.POINTER(SimSpec))
        check(lib().cave_serve_submit(self._h, ptr, request_id))

    def poll(self):
        out = []
        n = C.c_size_t(0)
        while True:
            check(lib().cave_serve_poll(self._h, self._buf, len(self._buf), C.byref(n)))
            out.extend(ServeCompletion.from_buffer_copy(self._buf[i]) for i in range(n.value))
            if n.value < len(self._buf):
                return out

    def drain(self):
        check(lib().cave_serve_drain(self._h))
        return self.poll()

    def stats(self) -> dict:
        v = [C.c_uint64() for _ in range(4)]
        check(lib().cave_serve_stats(self._h, *[C.byref(x) for x in v]))
        return dict(zip(("admitted", "finished", "resident_bytes", "peak_resident_bytes"), (x.value for x in v)))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().cave_serve_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
