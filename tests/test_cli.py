"""`cave` CLI argument handling (no device work): the reference's validation
messages and exit codes (src/main.cpp:76-84, 116-197, 343-377)."""
import os
import subprocess

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cave_amd", "_lib", "cave")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_help():
    r = run("--help")
    assert r.returncode == 0 and "--mode search" in r.stdout


def test_validation_errors_exit_1():
    assert run("--mode", "search", "--neighborhood", "X").returncode == 1
    assert run("--mode", "search", "--neighborhood", "W").returncode == 1  # W alone selects no neighbours
    assert run("--mode", "search", "--search-workgroup-size", "7").returncode == 1
    assert run("--mode", "search", "--spawn-mode", "sphere").returncode == 1
    assert run("--mode", "search", "--grid-sweep", "30:20:5").returncode == 1
    assert run("--mode", "search", "--shape", "dodecahedron").returncode == 1
    assert run("--mode", "video").returncode == 1
    r = run("--mode", "search", "--neighborhood", "X")
    assert "--neighborhood must include at least one of F / E / C" in r.stderr


def test_parse_errors_exit_2():
    assert run("--grid").returncode == 2
    assert run("--grid", "abc", "--mode", "search").returncode == 2
    assert run("positional").returncode == 2


def test_unknown_shape_exits_1():
    out = subprocess.run([CLI, "--mode", "search", "--shape", "hexagon"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 1 and "Unknown --shape" in (out.stdout + out.stderr)


def test_vram_budget_abort():
    """A chunk whose estimated VRAM (src/common/permutationMath.h:144-153)
    exceeds --gpu-vram-budget-gb aborts the sweep with the reference's error
    (src/modes/searchMode.cpp:358-366); the sweep then completes empty, as the
    reference's does, and the process exits 0.  512^3, FE, M 2..3: 6 slots x
    (3 x 4 B x 16,777,216 + 128 B) + 16 MiB = 1.14 GiB > 0.5 GiB.  No device
    work happens before the abort."""
    import tempfile

    with tempfile.TemporaryDirectory() as d:
        r = run("--mode", "search", "--grid", "512", "--neighborhood", "FE", "--min-cs", "2", "--max-cs-range", "3",
                "--gpu-vram-budget-gb", "0.5", "--output", d, "--headless")
        assert r.returncode == 0
        assert "Chunk requires 1.14 GB but VRAM budget is only 0.50 GB per GPU. Aborting sweep." in r.stderr
        assert "Sweep complete. Viable: 0, Possible: 0, Unviable: 0" in r.stdout
        # the budget below the 0.5 GB floor is raised to it (searchMode.cpp:227-229)
        r = run("--mode", "search", "--grid", "512", "--neighborhood", "FE", "--gpu-vram-budget-gb", "0.1",
                "--output", d, "--headless")
        assert "VRAM budget is only 0.50 GB" in r.stderr


def test_chunk_vram_estimate():
    """EstimateChunkVramBytes at hand-worked points: 49^3 holds 117,649 cells,
    14,707 packed words, so one slot takes 3 x 4 x 14,707 + 128 = 176,612 B;
    6 slots + 16,777,216 B of fixed overhead = 17,836,888 B."""
    import cave_amd as cv

    assert cv.estimate_chunk_vram_bytes(49, 49, 49, 6) == 17_836_888
    assert cv.estimate_chunk_vram_bytes(49, 49, 49, 0) == 16 << 20
    assert cv.estimate_chunk_vram_bytes(8, 1, 1, 1) == 12 + 128 + (16 << 20)
    assert cv.estimate_chunk_vram_bytes(9, 1, 1, 1) == 24 + 128 + (16 << 20)  # 9 cells round up to 2 words
    assert cv.estimate_chunk_vram_bytes(2048, 2048, 2048, 210) == (3 * 4 * 2**30 + 128) * 210 + (16 << 20)
