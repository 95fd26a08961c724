"""The register-tile planner on the CPU: tests/native/tile_reg_check.cpp runs
random op lists through reg_plan and the kernel's own per-thread steps and
compares with the documented op semantics (no GPU needed)."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


@pytest.mark.skipif(shutil.which(HIPCC) is None and not os.path.exists(HIPCC), reason="hipcc not available")
def test_register_tile_planner_matches_op_semantics(tmp_path):
    exe = tmp_path / "tile_reg_check"
    src = os.path.join(HERE, "native", "tile_reg_check.cpp")
    subprocess.run([HIPCC, "-std=c++17", "-O1", "--offload-arch=gfx950", src, "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    proc = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    lines = proc.stdout.strip().splitlines()
    assert proc.returncode == 0, proc.stdout[-2000:]
    assert len(lines) == 42 and all(line.startswith("ok ") for line in lines)
