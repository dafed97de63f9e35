"""The drop-in module loads and runs the way the reference harness drives the
reference solver (run_ab_case.py:261-279, 898-912): import from a file path
with the LBM_* environment set, module-level steps / plot_freq / visualize,
main(), then field reads with .to_numpy() and counters with [None]."""
import importlib.util
from pathlib import Path

import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of

pytestmark = pytest.mark.gpu

MODULE = Path(__file__).resolve().parents[1] / "lbm-3d-les-air-over-waves_amd" / "LBM_3D_LES_air_over_waves_hip.py"


def load_module(label):
    spec = importlib.util.spec_from_file_location(f"lbm_ab_{label}", MODULE)
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


def test_harness_style_run(monkeypatch):
    env = env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2, LBM_BED_WAVELENGTH_M=32)
    for key, value in env.items():
        monkeypatch.setenv(key, value)
    module = load_module("hip")
    assert (module.Nx, module.Ny, module.Nz) == (64, 16, 24)
    assert module.cx.to_numpy().shape == (19,)
    module.steps = 30
    module.plot_freq = 10
    module.plot_step_start = 0
    frames = []

    def visualize(step, time_phys, re_phys, re_lb):
        ux = module.ux.to_numpy()
        assert ux.shape == (module.Nz, module.Ny, module.Nx)
        frames.append((step, float(np.abs(ux).max()), int(module.boundary_cut_link_count[None])))

    module.visualize = visualize
    module.main()
    assert [f[0] for f in frames] == [10, 20, 30]
    assert all(np.isfinite(f[1]) and f[1] > 0 for f in frames)
    assert frames[-1][2] > 0  # the moving bed cuts links
    lo = module.lattice_open.to_numpy() if hasattr(module, "lattice_open") else None
    if lo is not None:
        assert lo.shape == (module.Nz, module.Ny, module.Nx, 19)
    assert module._solver.case.step_index() == 30
    module.close()
