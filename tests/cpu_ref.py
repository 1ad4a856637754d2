"""Helpers of the GPU tests that run the fast CPU form (oracle/cpu_fast.c,
held bit-exact to the scalar oracle by test_cpu_fast.py) on the same
simulation lists as the GPU."""
import os

import numpy as np

import oracle as O


def cpu_threads() -> int:
    """Host threads for CPU reference runs: the CPUs this process may use, at most 16."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def fast_sims(specs: np.ndarray) -> np.ndarray:
    """The fast CPU form's simulation records (oracle.SIM_NP) for a
    cave_amd.SIM_DTYPE array."""
    a = np.zeros(len(specs), dtype=O.SIM_NP)
    for k, f in enumerate(("X", "Y", "Z")):
        a[f] = specs["dims"][:, k]
    for k, f in enumerate(("sx", "sy", "sz")):
        a[f] = specs["spawn"][:, k]
    a["max_ticks"] = specs["ticks"]
    for f in ("M", "flags", "birth", "survival", "spawn_mode", "seed", "density"):
        a[f] = specs[f]
    return a


def run_cpu(specs: np.ndarray):
    """(ticks run, grid infos, final-cell hashes) of every simulation."""
    ticks, infos, ran, _, _, hashes = O.run_sims_fast(fast_sims(specs), n_threads=cpu_threads(), hashes=True)
    assert ran == len(specs)
    return ticks, infos, hashes


def assert_results_equal(res: np.ndarray, ticks, infos, hashes=None, label=""):
    """GPU RESULT_DTYPE rows against the CPU's, field by field."""
    names = {"died": "totalNumberOfCellsThatDied", "born": "totalNumberOfCellsThatWereBorn",
             "ticks_survived": "numberOfSimulationTicksSurvived", "status": "status",
             "max_alive": "maxAliveCount", "final_alive": "finalAliveCount"}
    assert np.array_equal(res["ticks_run"], ticks), f"{label} ticks run"
    for f, g in names.items():
        want = np.array([getattr(i, g) for i in infos], dtype=np.uint32)
        bad = np.nonzero(res[f] != want)[0]
        assert bad.size == 0, f"{label} {f}: first mismatch at {bad[:5]}"
    if hashes is not None:
        bad = np.nonzero(res["final_hash"] != hashes)[0]
        assert bad.size == 0, f"{label} final hash: first mismatch at {bad[:5]}"
