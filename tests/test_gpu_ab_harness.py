"""The metrics harness driving the drop-in module on the device: the rows it
writes equal the metrics of the oracle's state after the same steps
(run_ab_case.py:286-679 columns, run_ab_case.py:880-912 protocol)."""
import csv
import math
from pathlib import Path

import numpy as np
import pytest

import ab_harness as abh
import lbm_abi as abi
from harness import env_of, make_oracle
from lbm_solver import _COUNTERS, _FIELDS, _WATER_FIELDS, D3Q19_CX, D3Q19_CY, D3Q19_CZ, ConstField, CounterProxy, FieldProxy
from test_oracle_height import height_env

pytestmark = pytest.mark.gpu

MODULE = Path(__file__).resolve().parents[1] / "lbm-3d-les-air-over-waves_amd" / "LBM_3D_LES_air_over_waves_hip.py"

CASES = {
    "mound_height": height_env(LBM_FREE_SURFACE_INITIAL="gaussian", LBM_GAUSSIAN_AMPLITUDE_M=2.5,
                               LBM_GAUSSIAN_SIGMA_M=4, LBM_GAUSSIAN_CENTER_X_M=16),
    "air_waves": env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2,
                        LBM_BED_WAVELENGTH_M=32),
}

EXACT = {"label", "step", "cut_links", "fs_cut_links", "refill", "fs_refill", "fs_coalesce",
         "fs_bridge", "vof_excess_count", "vof_orphan_count", "vof_collapse_cells", "vof_tiny_interface_count",
         "vof_pending_excess_count", "vof_detached_count", "vof_compact_count"}


class OracleView:
    """The module surface over an oracle case, for the same metrics."""

    def __init__(self, cfg, case):
        self.case = case
        for name, value in cfg.values.items():
            setattr(self, name, value)
        self.cx, self.cy, self.cz = (ConstField(v, np.int32) for v in (D3Q19_CX, D3Q19_CY, D3Q19_CZ))
        dx = cfg.dx
        self.x_phys = (np.arange(self.Nx) + 0.5) * dx
        self.y_phys = (np.arange(self.Ny) + 0.5) * dx
        self.z_phys = (np.arange(self.Nz) + 0.5) * dx
        self.Z_phys, self.Y_phys, self.X_phys = np.meshgrid(self.z_phys, self.y_phys, self.x_phys, indexing="ij")
        self.FS_GAS, self.FS_LIQUID, self.FS_INTERFACE = abi.FS_GAS, abi.FS_LIQUID, abi.FS_INTERFACE
        fields = dict(_FIELDS)
        if cfg.water_free_surface_enabled:
            fields.update(_WATER_FIELDS)
        for name, fid in fields.items():
            setattr(self, name, FieldProxy(self, fid))
        for name in _COUNTERS:
            setattr(self, name, CounterProxy(self, name))

    def counters(self):
        return self.case.counters()


# the fastest cell and what is around it: the argmax can move between
# cells whose speeds agree to rounding
LOCATION = {"max_speed_x", "max_speed_y", "max_speed_z", "max_speed_fs_type", "max_speed_fs_fill",
            "max_speed_wet_neighbors", "max_speed_full_neighbors", "max_speed_gas_neighbors"}


def tolerance(key, view):
    """FP32 device against the FP32 oracle: fields agree to a few ulp, so
    level columns get 1e-4 relative; rms columns of differences and
    residuals lose digits and get 2e-2 relative plus an absolute floor of
    the rounding of their inputs (rho to 2e-6, u to 2e-6, FP32 heights of
    order 10 m to 1e-5 m)."""
    if key.endswith("rms") or "resid" in key or key.startswith("surface_symmetry"):
        if "_p_" in key or key.startswith("p_"):
            return 2e-2, view.rho_phys * view.vel_scale**2 * 2e-6 / 3.0
        if "vort" in key:
            return 2e-2, 2e-6
        return 2e-2, 1e-5
    return 1e-4, 0.0


def close(a, b, rel, atol):
    if isinstance(b, float) and math.isnan(b):
        return math.isnan(a)
    return abs(a - b) <= rel * abs(b) + atol or abs(a - b) <= 1e-30


@pytest.mark.parametrize("name", list(CASES))
def test_harness_rows_match_oracle_metrics(name, tmp_path, monkeypatch):
    env = CASES[name]
    for key, value in env.items():
        monkeypatch.setenv(key, value)
    print(f"harness case {name}", flush=True)
    out = abh.run(MODULE, name, steps=40, plot_freq=20, plot_start=0, output_dir=tmp_path)
    with out.open() as fh:
        rows = list(csv.DictReader(fh))
    assert [int(r["step"]) for r in rows] == [20, 40]
    cfg, o = make_oracle(env)
    view = OracleView(cfg, o)
    done = 0
    for r in rows:
        step = int(r["step"])
        o.step(step - done)
        done = step
        ref = abh.frame_metrics(view, name, step, (step + 1) * cfg.dt_phys)
        for key in abh.METRIC_FIELDNAMES:
            got = r[key]
            want = ref[key]
            if key in EXACT:
                assert got == str(want), (name, step, key, got, want)
            elif key not in LOCATION:
                g, w = float(got), float(want)
                assert close(g, w, *tolerance(key, view)), (name, step, key, g, w)
    assert float(rows[-1]["max_speed_phys"]) > 0.0
    if cfg.water_free_surface_enabled:
        assert float(rows[-1]["water_volume"]) > 0.0 and math.isfinite(float(rows[-1]["p_linear_resid_rms"]))
    else:
        assert int(rows[-1]["cut_links"]) > 0
    o.close()
