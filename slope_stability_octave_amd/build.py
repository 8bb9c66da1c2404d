"""Build libssr_hip.so in-tree with hipcc for gfx950.

All HIP sources under csrc/ are compiled into one shared library next to this
file.  Flags: -O3, gfx950 only, packed-FP32 instructions disabled (the compiler
otherwise forms them from plain float code); the host drivers (*.cpp) are
built with OpenMP for the threaded AMG setup.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libssr_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-munsafe-fp-atomics",
]


def sources():
    return sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(CSRC, os.path.basename(src) + ".o")
    cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-c", src, "-o", obj, "-I", CSRC]
    if src.endswith(".cpp"):
        # host-side drivers (AMG setup uses OpenMP threads)
        cmd = [HIPCC] + [f for f in FLAGS if f != "-shared"] + ["-fopenmp", "-x", "hip", "-c", src, "-o", obj,
                                                               "-I", CSRC]
    if verbose:
        print(" ".join(cmd), file=sys.stderr)
    subprocess.run(cmd, check=True)
    return obj


def build(verbose: bool = False) -> str:
    """Compile every source into libssr_hip.so (always recompiles; the
    translation units are compiled concurrently)."""
    from concurrent.futures import ThreadPoolExecutor

    srcs = sources()
    jobs = max(1, min(len(srcs), os.cpu_count() or 1, 8))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
