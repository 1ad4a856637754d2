"""Build recipe of the native library (libcave_amd.so) and the `cave` CLI.

Every HIP source is compiled ahead of time for gfx950 with hipcc; outputs go
to cave_amd/_lib/ (in-tree, git-ignored) so that they travel with the tree.

    python -m cave_amd.build            # build (incremental)
    python -m cave_amd.build --force    # rebuild everything
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_lib")
OBJ = os.path.join(OUT, "obj")
LIB = os.path.join(OUT, "libcave_amd.so")
CLI = os.path.join(OUT, "cave")

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
COMMON = ["-O3", "-std=c++20", "-fPIC", "-Wall", "-Wno-unused-function", f"-I{ROOT}/include", f"-I{CSRC}"]
# packed-FP32 instructions stay disabled in every device compile
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]

LIB_SOURCES = [
    "api_common.cpp",
    "api_step.cpp",
    "api_batch.cpp",
    "api_serve.cpp",
    "api_search.cpp",
    "api_aggregate.cpp",
    "batch_plan.cpp",
    "batch.hip",
    "step_gather.hip",
    "step_stream.hip",
    "step_erd.hip",
    "api_static.hip",
]
CLI_SOURCES: list[str] = ["cli.cpp"]


def _headers_mtime() -> float:
    ts = [os.path.getmtime(os.path.join(ROOT, "include", "cave.h"))]
    for f in os.listdir(CSRC):
        if f.endswith(".h"):
            ts.append(os.path.getmtime(os.path.join(CSRC, f)))
    return max(ts)


def _compile(src: str, force: bool) -> str:
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, src + ".o")
    if not force and os.path.exists(obj):
        if os.path.getmtime(obj) >= max(os.path.getmtime(path), _headers_mtime()):
            return obj
    lang = ["-x", "hip"] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [HIPCC, *COMMON, *DEVICE, *lang, "-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(force: bool = False, jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), LIB_SOURCES))
    newest = max(os.path.getmtime(o) for o in objs)
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < newest:
        tmp = LIB + ".tmp"
        cmd = [HIPCC, *DEVICE, "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
            raise RuntimeError("link of libcave_amd.so failed")
        os.replace(tmp, LIB)
    if CLI_SOURCES:
        cobjs = [_compile(s, force) for s in CLI_SOURCES]
        if force or not os.path.exists(CLI) or os.path.getmtime(CLI) < max(os.path.getmtime(LIB), *map(os.path.getmtime, cobjs)):
            cmd = [HIPCC, *DEVICE, "-o", CLI, *cobjs, f"-L{OUT}", "-lcave_amd", "-Wl,-rpath,$ORIGIN", "-lpthread"]
            r = subprocess.run(cmd, capture_output=True, text=True)
            if r.returncode != 0:
                sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
                raise RuntimeError("link of the cave CLI failed")
    return LIB


def clean() -> None:
    shutil.rmtree(OUT, ignore_errors=True)


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build(force=a.force, jobs=a.j))
