"""Builds ``libqjhip.so`` from ``csrc/*.hip`` for gfx950 with hipcc.

Every translation unit is compiled in parallel, then linked into one shared
library next to this file.  ``python -m qibojit_amd.build`` runs it.
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqjhip.so")
ARCH = os.environ.get("QJ_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["qj_runtime.hip", "qj_gates.hip", "qj_ops.hip", "qj_tile.hip", "qj_group.hip", "qj_mfma.hip",
           "qj_clifford.hip"]

CFLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # packed FP32 instructions are disabled: they are generated from plain
    # float code otherwise, and some of their forms are not trusted here.
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-result",
]


# Per-source flags.  Branches on wave-uniform values (the op dispatch of the
# register tile kernel, the per-step rotation forms of the rotation passes)
# stay unstructured, which keeps register-resident amplitudes from being
# copied around every branch.
_UNIFORM_CFG = ["-mllvm", "-structurizecfg-skip-uniform-regions=true"]
EXTRA = {"qj_tile.hip": _UNIFORM_CFG, "qj_group.hip": _UNIFORM_CFG, "qj_clifford.hip": _UNIFORM_CFG}


def _run(cmd):
    proc = subprocess.run(cmd, capture_output=True, text=True, check=False)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{proc.stdout}\n{proc.stderr}")
    return proc


def build(verbose=False):
    """Compile every HIP source and link ``libqjhip.so``; returns its path."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    objs = []

    def compile_one(src):
        obj = os.path.join(BUILD_DIR, src.replace(".hip", ".o"))
        cmd = [HIPCC, *CFLAGS, *EXTRA.get(src, []), "-c", os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as pool:
        objs = list(pool.map(compile_one, SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-Wl,-rpath,/opt/rocm/lib"]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
