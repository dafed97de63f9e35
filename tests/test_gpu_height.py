"""Parity of the device height-function free-surface tracking (staged path,
CODEX:1209-1232, 2163-2316, 2376-2429) with the oracle: a Gaussian mound in
a periodic box, the reference's default water set-up (solitary wave, solid
x walls, hydrostatic-balanced pressure) and the total-pressure variant.

Also the surface prescribed as the travelling solitary wave and the
free-surface population projection at initialisation (CODEX:2829-2875), in
height and VOF tracking.

Bar: cell types bit-exact; column heights, fill, populations, velocity and
density within max-relative 1e-5 after 60 steps; refill counters equal."""
import numpy as np
import pytest

import lbm_abi as abi
from test_gpu_vof import pair, relerr
from test_oracle_height import height_env

pytestmark = pytest.mark.gpu

CASES = {
    "mound_periodic": height_env(LBM_FREE_SURFACE_INITIAL="gaussian", LBM_GAUSSIAN_AMPLITUDE_M=2.5,
                                 LBM_GAUSSIAN_SIGMA_M=4, LBM_GAUSSIAN_CENTER_X_M=16),
    "solitary_solid_x": height_env(LBM_FREE_SURFACE_INITIAL="solitary", LBM_SOLITARY_AMPLITUDE_M=3,
                                   LBM_X_BOUNDARY="solid", LBM_SOLITARY_X0_M=10, LBM_COLLISION="kbc"),
    "mound_total_pressure": height_env(LBM_FREE_SURFACE_INITIAL="gaussian", LBM_GAUSSIAN_AMPLITUDE_M=2.5,
                                       LBM_GAUSSIAN_SIGMA_M=4, LBM_PRESSURE_FORMULATION="total_pressure",
                                       LBM_X_BOUNDARY="solid"),
    "prescribed_solitary": height_env(LBM_FREE_SURFACE_MODE="prescribed_solitary", LBM_FREE_SURFACE_INITIAL="solitary",
                                      LBM_SOLITARY_AMPLITUDE_M=2.0, LBM_SOLITARY_X0_M=8.0),
    "mound_projected_init": height_env(LBM_FREE_SURFACE_INITIAL="gaussian", LBM_GAUSSIAN_AMPLITUDE_M=2.5,
                                       LBM_GAUSSIAN_SIGMA_M=4, LBM_INIT_FREE_SURFACE_POPULATIONS=1),
}


@pytest.mark.parametrize("name", list(CASES))
def test_height_tracking_parity(name):
    cfg, o, c = pair(CASES[name])
    assert np.array_equal(c.download(abi.FIELD_FS_TYPE), o.download(abi.FIELD_FS_TYPE))
    h0 = o.download(abi.FIELD_FS_H).copy()
    for chunk in (20, 20, 20):
        o.step(chunk)
        c.step(chunk)
        assert np.array_equal(c.download(abi.FIELD_FS_TYPE), o.download(abi.FIELD_FS_TYPE)), (name, c.step_index())
        for fid in (abi.FIELD_FS_H, abi.FIELD_FS_FILL, abi.FIELD_F, abi.FIELD_UX, abi.FIELD_RHO):
            assert relerr(c.download(fid), o.download(fid)) <= 1e-5, (name, c.step_index(), fid)
    # the surface moved
    assert np.max(np.abs(o.download(abi.FIELD_FS_H) - h0)) > 1e-6
    cc, oc = c.counters(), o.counters()
    for key in ("free_surface_refill_count", "free_surface_cut_link_count", "boundary_cut_link_count"):
        assert cc[key] == oc[key], key
    c.close()
    o.close()


def test_vof_projected_init_parity():
    from test_oracle_vof import vof_env

    cfg, o, c = pair(vof_env(LBM_INIT_FREE_SURFACE_POPULATIONS=1))
    assert np.array_equal(c.download(abi.FIELD_F), o.download(abi.FIELD_F))
    o.step(40)
    c.step(40)
    assert np.array_equal(c.download(abi.FIELD_FS_TYPE), o.download(abi.FIELD_FS_TYPE))
    for fid in (abi.FIELD_FS_FILL, abi.FIELD_FS_MASS, abi.FIELD_F, abi.FIELD_UZ):
        assert relerr(c.download(fid), o.download(fid)) <= 1e-6, fid
    c.close()
    o.close()
