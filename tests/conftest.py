"""Test configuration: the `gpu` marker and import paths.

CPU tests (-m "not gpu") pin the CPU oracle against closed forms and the
host logic; GPU tests (-m gpu) run liblbm.so on the device and compare it
with the oracle.
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "lbm-3d-les-air-over-waves_amd"
for p in (str(PKG), str(ROOT / "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and liblbm.so")


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle_ref import load_oracle

    return load_oracle()
