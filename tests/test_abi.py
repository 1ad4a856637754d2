"""The C-ABI library loads on the CPU host and exports every declared symbol.

No compute call is made here (there is no GPU); host-only helpers that need no
device are checked against the oracle's restatements.
"""
import ctypes as C
import os
import re

import pytest

import cave_amd
import oracle as O


def declared_functions():
    text = open(cave_amd.HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cave_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(cave_amd.LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python mirror binds exactly the declared set
    assert sorted(cave_amd._SIGS) == names


def test_struct_sizes_match_reference_layouts():
    # structs.h:79-136 sizes: SimulationParameters 24, GridSnapshot 16, GridInfo 24, ChunkConfiguration 56
    assert C.sizeof(cave_amd.SimulationParameters) == 24
    assert C.sizeof(cave_amd.GridSnapshot) == 16
    assert C.sizeof(cave_amd.GridInfo) == 24
    assert C.sizeof(cave_amd.ChunkConfiguration) == 56
    assert C.sizeof(cave_amd.TickStats) == 48
    # this library's own records (no reference counterpart)
    assert C.sizeof(cave_amd.SimSpec) == 72
    assert C.sizeof(cave_amd.SimResult) == 88
    assert C.sizeof(cave_amd.BatchOptions) == 40
    assert cave_amd.TIMING_DTYPE.itemsize == 32
    assert C.sizeof(cave_amd.ServeOptions) == 24
    assert C.sizeof(cave_amd.ServeCompletion) == 136


def test_abi_version():
    assert cave_amd.lib().cave_abi_version() == 2


def test_host_rule_strings_match_oracle():
    for s in ["3,5-7", "2,4", "1-26", " 4 , 9-11,60,61", "0", "7-5", "1,2", "1,2,3,5"]:
        m = cave_amd.rule_string_to_bitmask(s)
        assert m == O.rule_string_to_bitmask(s)
        assert cave_amd.bitmask_to_rule_string(m) == O.bitmask_to_rule_string(m)


@pytest.mark.parametrize("N,K", [(6, 2), (12, 2), (18, 2), (26, 2), (26, 4), (7, 7), (18, 0)])
def test_host_enumeration_matches_oracle(N, K):
    assert cave_amd.enumerate_valid_bitmasks(N, K) == O.enumerate_valid_bitmasks(N, K)


def test_cell_words():
    assert cave_amd.cell_words(49, 49, 49) == 7 * 49 * 49 == O.cell_words(49, 49, 49)
    assert cave_amd.cell_words(512, 512, 512) == 64 * 512 * 512


def test_splitmix64_known_values():
    # splitmix64 reference outputs for seed 0 (Vigna's published generator)
    r = cave_amd.Rng(0)
    assert r.next() == 0xE220A8397B1DCDAF
    assert r.next() == 0x6E789E6AA1B965F4
    assert r.next() == 0x06C45D188009454F


def _sim(**kw):
    base = dict(dims=(8, 8, 8), M=2, flags=cave_amd.FACES, ticks=4, birth=0b100, survival=0b1100, spawn=(4, 4, 4))
    base.update(kw)
    return cave_amd.Sim(**base)


@pytest.mark.parametrize("kw,needle", [
    (dict(dims=(0, 8, 8)), "dims"),
    (dict(dims=(8, 4097, 8)), "dims"),
    (dict(dims=(4096, 4096, 4096)), "2^32"),
    (dict(M=0), "max_cell_state"),
    (dict(M=17), "max_cell_state"),
    (dict(flags=16), "neighborhood"),
    (dict(ticks=0), "ticks"),
    (dict(spawn_mode=cave_amd.SPAWN_BUFFER), "spawn_mode"),
    (dict(sim_flags=1 << 20), "sim flags"),
])
def test_batch_rejects_bad_specs_before_device_work(kw, needle):
    # every spec is validated before the first HIP call, so this runs without
    # a GPU; the error names the offending spec
    with pytest.raises(cave_amd.CaveError) as e:
        cave_amd.batch_run([_sim(), _sim(**kw)])
    assert e.value.code == cave_amd.CAVE_E_INVALID
    assert "spec 1" in str(e.value) and needle in str(e.value)


def test_batch_of_nothing_is_ok():
    # an empty batch returns at once, before any device work
    res, rows, _ = cave_amd.batch_run([])
    assert len(res) == 0 and rows is None


@pytest.mark.parametrize("kw,needle", [
    (dict(dims=(0, 4, 4)), "grid dimensions"),
    (dict(dims=(65537, 1, 1)), "grid dimensions"),
    # cube grids may exceed 2^32 cells (their u32 counters wrap, as the
    # reference's do); the ERD kernel indexes cells with 32 bits
    (dict(dims=(4096, 4096, 4096), shape=cave_amd.SHAPE_ERD, kernel=cave_amd.KERNEL_ERD), "2^32"),
    (dict(spawn_mode=99), "spawn_mode"),
    (dict(spawn_mode=cave_amd.SPAWN_BUFFER), "initial_cells"),
    (dict(kernel=99), "kernel selector"),
])
def test_step_rejects_bad_specs_before_device_work(kw, needle):
    args = dict(dims=(8, 8, 8))
    args.update(kw)
    with pytest.raises(cave_amd.CaveError) as e:
        cave_amd.Step(args.pop("dims"), 2, cave_amd.FACES, 0b100, 0b1100, **args)
    assert e.value.code == cave_amd.CAVE_E_INVALID
    assert needle in str(e.value)


@pytest.mark.parametrize("kw,code,needle", [
    (dict(flags=0), cave_amd.CAVE_E_INVALID, "must include F, E or C"),
    (dict(flags=16), cave_amd.CAVE_E_INVALID, "bad neighbourhood flags"),
    (dict(grid=0), cave_amd.CAVE_E_INVALID, "grid dimensions"),
    (dict(max_cs=40), cave_amd.CAVE_E_INVALID, "max cell state range"),
    (dict(ticks=0), cave_amd.CAVE_E_INVALID, "ticks"),
    (dict(density_index=16), cave_amd.CAVE_E_INVALID, "density index"),
    (dict(partition_index=3, partition_count=3), cave_amd.CAVE_E_INVALID, "partition index"),
    (dict(shape=5), cave_amd.CAVE_E_UNSUPPORTED, "shape ids"),
])
def test_search_rejects_bad_specs_before_device_work(kw, code, needle):
    args = dict(grid=9, flags=cave_amd.FACES, max_rule_bits=1, min_cs=2, max_cs=2, ticks=4)
    args.update(kw)
    with pytest.raises(cave_amd.CaveError) as e:
        cave_amd.Search(args.pop("grid"), args.pop("flags"), args.pop("max_rule_bits"), args.pop("min_cs"),
                        args.pop("max_cs"), args.pop("ticks"), **args)
    assert e.value.code == code
    assert needle in str(e.value)
