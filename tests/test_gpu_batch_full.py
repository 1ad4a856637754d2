"""The batched search workloads in full on the GPU against the fast CPU form,
simulation by simulation: every grid-info field, the ticks run and the final
cells' hash.  The uniform search: 4096 rule configs x 32 seeds at 49^3, FE,
C in 2..5, 20 ticks (131,072 simulations, one persistent launch, LDS-resident
grids).  The uneven search: 2048 configs, edge 24..160, F/FE/FEV, C in
2..16, 20..200 ticks, autotuned and split over blocks.  Reference semantics:
searchSimulationCube.slang:78-247, searchManager.slang:98-149."""
import numpy as np
import pytest

import cave_amd as cv
from cave_amd import workloads as W
from cpu_ref import assert_results_equal, run_cpu

pytestmark = pytest.mark.gpu


def test_uniform_49_FE_search_every_simulation():
    # final-cell hashes on every 16th simulation (the engine stages the final
    # cells of each hashed simulation: 67 KB at 49^3)
    specs = W.search_uniform(4096, 32)
    specs["sim_flags"][::16] = cv.SIM_FINAL_HASH
    print(f"uniform 49^3 search: {len(specs)} simulations on the GPU", flush=True)
    res, _ = cv.batch_run_np(specs, engine=cv.ENGINE_PERSISTENT)
    print("uniform 49^3 search: the same list on the fast CPU form", flush=True)
    ticks, infos, hashes = run_cpu(specs)
    assert_results_equal(res, ticks, infos, None, "uniform 49^3 search")
    assert np.array_equal(res["final_hash"][::16], hashes[::16]), "uniform 49^3 search final hash"


def test_uneven_24_160_search_every_simulation():
    specs = W.search_uneven(2048)
    specs["sim_flags"] = cv.SIM_FINAL_HASH
    print(f"uneven search: {len(specs)} simulations on the GPU (autotuned, split)", flush=True)
    res, _, tm, _ = cv.batch_run_np(specs, timing=True)
    assert (tm["blocks"] > 1).sum() > 100
    print("uneven search: the same list on the fast CPU form", flush=True)
    ticks, infos, hashes = run_cpu(specs)
    assert int(ticks.max()) > 150
    assert_results_equal(res, ticks, infos, hashes, "uneven search")
