"""The multiply-high cell-index division of the staged kernels equals
integer division for every cell index a grid can have (host build of the
same inline functions; no GPU)."""
import shutil
import subprocess
from pathlib import Path

import pytest

HERE = Path(__file__).resolve().parent


@pytest.mark.skipif(shutil.which("hipcc") is None and not Path("/opt/rocm/bin/hipcc").exists(), reason="hipcc not found")
def test_fast_div_equals_integer_division(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    exe = tmp_path / "fast_div_check"
    src = HERE / "native" / "fast_div_check.hip"
    build = subprocess.run([hipcc, "--offload-arch=gfx950", "-O2", "-std=c++17", "-Wno-unused-command-line-argument", str(src), "-o",
                            str(exe)], capture_output=True, text=True, timeout=600)
    assert build.returncode == 0, build.stderr[-3000:]
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600)
    assert run.returncode == 0, run.stdout + run.stderr
    assert "bad=0" in run.stdout
