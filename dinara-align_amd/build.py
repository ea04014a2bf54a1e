"""Builds libdinara_hip.so, the MI355X library, in-tree with hipcc for gfx950 (and probe/libdinara_probe.so, the
roofline probes bench.py measures its denominators with).

Each HIP source under csrc/ compiles to an object in parallel, then one shared library links them, placed in
the Python package folder beside the ctypes binding that loads it. Always compiles: a build is the check that
the sources compile.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "_build")
LIBRARY = os.path.join(HERE, "dinara_align", "libdinara_hip.so")
PROBE_SOURCE = os.path.join(HERE, "probe", "probe.hip")
PROBE = os.path.join(HERE, "probe", "libdinara_probe.so")
ROOT = os.path.dirname(HERE)

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-I" + os.path.join(ROOT, "include"),
    # Packed FP32 forms are kept out of every device code object.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def sources() -> list:
    return sorted(os.path.join(CSRC, name) for name in os.listdir(CSRC) if name.endswith(".hip"))


def compile_one(source: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.basename(source).replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", source, "-o", obj]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(result.stdout + result.stderr)
        raise RuntimeError(f"hipcc failed on {source}")
    return obj


def build(jobs: int = 8) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        objects = list(pool.map(compile_one, sources()))
    tmp = LIBRARY + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objects, "-lpthread"]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(result.stdout + result.stderr)
        raise RuntimeError("linking libdinara_hip.so failed")
    os.replace(tmp, LIBRARY)
    # The roofline probes (bench.py only): a library of their own.
    tmp = PROBE + ".tmp"
    cmd = [HIPCC, *FLAGS, "-shared", "-o", tmp, PROBE_SOURCE]
    result = subprocess.run(cmd, capture_output=True, text=True)
    if result.returncode != 0:
        sys.stderr.write(result.stdout + result.stderr)
        raise RuntimeError("building libdinara_probe.so failed")
    os.replace(tmp, PROBE)
    return LIBRARY


if __name__ == "__main__":
    print(build())
