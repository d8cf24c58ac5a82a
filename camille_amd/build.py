"""Builds the HIP C-ABI library camille_amd/liblee_amd.so for gfx950.

    python -m camille_amd.build

Compiles every source of camille_amd/csrc with hipcc into one shared
library; nothing is JIT-compiled at import time.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "liblee_amd.so")
SOURCES = ["lee_api.hip", "lee_kernels.hip", "lee_kernels3d.hip", "lee_adjoint.hip", "lee_adjoint3d.hip", "lee_noalloc.hip", "lee_plan.cpp",
           "lee_io.cpp"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
OBJDIR = os.path.join(HERE, "_build")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]


def build(verbose=False):
    """Compiles the sources in parallel (one hipcc per file), then links."""
    os.makedirs(OBJDIR, exist_ok=True)
    objs = [os.path.join(OBJDIR, s + ".o") for s in SOURCES]

    def compile_one(k):
        cmd = [HIPCC] + FLAGS + ["-c", os.path.join(CSRC, SOURCES[k]), "-o", objs[k]]
        if verbose:
            print(" ".join(cmd), flush=True)
        return subprocess.run(cmd, capture_output=True, text=True)

    with ThreadPoolExecutor(len(SOURCES)) as ex:
        results = list(ex.map(compile_one, range(len(SOURCES))))
    for k, r in enumerate(results):
        if r.stdout or r.stderr:
            sys.stderr.write(r.stdout + r.stderr)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {SOURCES[k]}")
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
