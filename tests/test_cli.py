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
