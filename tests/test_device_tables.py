"""The device-side D3Q19 tables (constexpr functions in
lbm-3d-les-air-over-waves_amd/csrc/lbm_device.h) against the reference's
tables (lbm_lattice.py:3-43) and the geometric definitions of the opposite,
z-mirror and column-sample maps.  Host build of the header with hipcc."""
import shutil
import subprocess
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
HDR = ROOT / "lbm-3d-les-air-over-waves_amd" / "csrc" / "lbm_device.h"
CX = [0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0]
CY = [0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, -1, 1]
CZ = [0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1]
W = [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12


def test_device_tables_match_reference():
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(hipcc).exists():
        pytest.skip("hipcc not available")
    src = f'''#include "{HDR}"
#include <cstdio>
int main() {{
  for (int q = 0; q < 19; q++)
    std::printf("%d %d %d %d %d %d %.17g\\n", lbm::cxq(q), lbm::cyq(q), lbm::czq(q), lbm::oppq(q), lbm::mirzq(q),
                lbm::c9q(q), lbm::wq<double>(q));
  return 0;
}}'''
    with tempfile.TemporaryDirectory() as d:
        cpp = Path(d) / "t.hip"
        cpp.write_text(src)
        exe = Path(d) / "t"
        subprocess.run([hipcc, "--offload-arch=gfx950", str(cpp), "-o", str(exe)], check=True, capture_output=True)
        rows = [list(map(float, l.split())) for l in subprocess.run([str(exe)], capture_output=True, text=True,
                                                                     check=True).stdout.strip().split("\n")]
    for q, (cx, cy, cz, opp, mir, c9, w) in enumerate(rows):
        assert (cx, cy, cz) == (CX[q], CY[q], CZ[q]), q
        assert (CX[int(opp)], CY[int(opp)], CZ[int(opp)]) == (-CX[q], -CY[q], -CZ[q]), q
        assert (CX[int(mir)], CY[int(mir)], CZ[int(mir)]) == (CX[q], CY[q], -CZ[q]), q
        assert c9 == (CX[q] + 1) * 3 + (CY[q] + 1)
        assert w == W[q]
