"""Parity of the box obstacle and of the phi boundary-geometry modes on the
device (staged path) with the oracle: geometry fields bit-exact, then 100
steps within the FP32 bar (max-relative 1e-5) for a box above a moving sine
bed, a yawed and pitched box in a channel with inflow, the moving bed with
fractions and flags from phi, and the box over the bed with fractions from
phi."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of
from test_gpu_parity import F32_TOL, GEOM_FIELDS, device_case, oracle_case, relerr

pytestmark = pytest.mark.gpu

CASES = {
    "bed_cube_waves": env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2,
                             LBM_BED_WAVELENGTH_M=32, LBM_OBSTACLE_MODE="bed_cube", LBM_CUBE_SIDE_M=6,
                             LBM_CUBE_CENTER_X_M=30.4, LBM_CUBE_CENTER_Y_M=8.1, LBM_CUBE_CENTER_Z_M=13.3),
    "cube_yaw_pitch": env_of("plain_64x32x32", LBM_LX_M=48, LBM_LY_M=16, LBM_LZ_M=24, LBM_OBSTACLE_MODE="cube",
                             LBM_CUBE_SIZE_X_M=8, LBM_CUBE_SIZE_Y_M=5, LBM_CUBE_SIZE_Z_M=6, LBM_CUBE_YAW_DEG=25,
                             LBM_CUBE_PITCH_DEG=15, LBM_CUBE_CENTER_X_M=20.2, LBM_CUBE_CENTER_Y_M=7.7,
                             LBM_CUBE_CENTER_Z_M=11.1, LBM_COLLISION="kbc"),
    "phi_waves": env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2,
                        LBM_BED_WAVELENGTH_M=32, LBM_BOUNDARY_GEOMETRY="phi"),
    "phi_fraction_bed_cube": env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2,
                                    LBM_BED_WAVELENGTH_M=32, LBM_OBSTACLE_MODE="bed_cube", LBM_CUBE_SIDE_M=6,
                                    LBM_CUBE_CENTER_X_M=30.4, LBM_CUBE_CENTER_Y_M=8.1, LBM_CUBE_CENTER_Z_M=13.3,
                                    LBM_BOUNDARY_GEOMETRY="phi_fraction"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_box_geometry_bit_exact(name):
    env = CASES[name]
    _, o = oracle_case(env)
    _, c = device_case(env)
    for fid in GEOM_FIELDS:
        a, b = c.download(fid), o.download(fid)
        assert np.array_equal(a, b), (name, fid, int(np.sum(a != b)))
    c.close()
    o.close()


@pytest.mark.parametrize("name", list(CASES))
def test_box_step_parity(name):
    env = CASES[name]
    _, o = oracle_case(env)
    _, c = device_case(env)
    assert c.path() == abi.PATH_STAGED
    o.step(100)
    c.step(100)
    for fid in (abi.FIELD_F, abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UZ, abi.FIELD_OMEGA):
        assert relerr(c.download(fid), o.download(fid)) <= F32_TOL, (name, fid)
    assert c.counters()["boundary_cut_link_count"] == o.counters()["boundary_cut_link_count"]
    c.close()
    o.close()
