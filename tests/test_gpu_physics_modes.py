"""Water without a free surface and air with one, on the device against the
oracle: the physics mode decides the static-wall conditioning
(CODEX:2922-2940), the free surface every free-surface rule (CODEX:49-60).
Both run on the staged path."""
import numpy as np
import pytest

import lbm_abi as abi
from test_gpu_vof import pair, relerr
from test_oracle_physics_modes import channel_env

pytestmark = pytest.mark.gpu

CASES = {
    "water_static_bed": channel_env("water", LBM_USE_LES=1, LBM_INITIAL_STATE="uniform_noise", LBM_INIT_U_LU=0.04,
                                    LBM_INIT_NOISE_LU=1e-3),
    "water_moving_bed": channel_env("water", LBM_BED_PROFILE="moving_sine", LBM_BED_AMP_M=1.0, LBM_BED_WAVELENGTH_M=16,
                                    LBM_COLLISION="kbc", LBM_USE_LES=1),
    "air_height_surface": channel_env("air", LBM_FREE_SURFACE_MODE="height_kinematic", LBM_FREE_SURFACE_DEPTH_M=6.4,
                                      LBM_FREE_SURFACE_INITIAL="gaussian", LBM_GAUSSIAN_AMPLITUDE_M=1.5,
                                      LBM_GAUSSIAN_SIGMA_M=3, LBM_GAUSSIAN_CENTER_X_M=8, LBM_G_LB="1e-4",
                                      LBM_PRESSURE_FORMULATION="hydrostatic_balanced"),
}


@pytest.mark.parametrize("name", list(CASES))
def test_physics_mode_parity(name):
    print(f"physics-mode case {name}", flush=True)
    cfg, o, c = pair(CASES[name])
    assert c.path() == abi.PATH_STAGED
    fids = [abi.FIELD_F, abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UZ]
    if cfg.water_free_surface_enabled:
        fids += [abi.FIELD_FS_H, abi.FIELD_FS_FILL]
    for chunk in (20, 20, 20):
        o.step(chunk)
        c.step(chunk)
        for fid in fids:
            assert relerr(c.download(fid), o.download(fid)) <= 1e-5, (name, c.step_index(), fid)
        if cfg.water_free_surface_enabled:
            assert np.array_equal(c.download(abi.FIELD_FS_TYPE), o.download(abi.FIELD_FS_TYPE)), name
    cc, oc = c.counters(), o.counters()
    for key in ("boundary_cut_link_count", "boundary_refill_count", "free_surface_refill_count"):
        assert cc[key] == oc[key], (name, key)
    assert np.max(np.abs(c.download(abi.FIELD_UX))) > 0.0
    c.close()
    o.close()


def test_link_counters_describe_the_last_step():
    """The stream's link counters are counted on the last step of a call:
    after step(5) they equal those after five step(1) calls, and those of
    the oracle after the same steps (which counts every step)."""
    env = CASES["air_height_surface"]
    cfg, o, a = pair(env)
    from harness import setup_case
    from lbm_native import Case

    b = Case(cfg.to_struct())
    setup_case(b, cfg)
    a.step(5)
    for _ in range(5):
        b.step(1)
    o.step(5)
    keys = ("boundary_cut_link_count", "free_surface_cut_link_count", "free_surface_thin_gap_bridge_count")
    ca, cb, co = a.counters(), b.counters(), o.counters()
    for key in keys:
        assert ca[key] == cb[key] == co[key], (key, ca[key], cb[key], co[key])
    assert ca["free_surface_cut_link_count"] > 0
    assert np.array_equal(a.download(abi.FIELD_F), b.download(abi.FIELD_F))
    for c in (a, b, o):
        c.close()
