"""Physics mode and free surface are independent switches in the reference
(CODEX:49-60: water_free_surface_enabled follows LBM_FREE_SURFACE_MODE
only).  The physics mode alone decides whether the closed links of static
walls are reset to the rest equilibrium each step (the air branch of
update_bed_populations_and_reinit, CODEX:2922-2940); the free surface alone
decides every free-surface rule.  Pinned on the oracle with closed forms."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import make_oracle
from lbm_config import parse_env

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
STAGE_BED_POPULATIONS = 4


def channel_env(physics, **kw):
    env = {
        "LBM_PHYSICS_MODE": physics, "LBM_FREE_SURFACE_MODE": "none", "LBM_COLLISION": "regularized",
        "LBM_USE_LES": "0", "LBM_TAU": "0.6", "LBM_DX_M": "1", "LBM_LX_M": "16", "LBM_LY_M": "8", "LBM_LZ_M": "12",
        "LBM_BED_PROFILE": "flat", "LBM_BED_LEVEL_M": "2.3", "LBM_X_BOUNDARY": "periodic",
        "LBM_ENABLE_TOP_DRIVE": "0", "LBM_ENABLE_X_FRINGE": "0", "LBM_ENABLE_TOP_SPONGE": "0", "LBM_BACKSCATTER": "0",
    }
    env.update({k: str(v) for k, v in kw.items()})
    return env


def test_water_without_a_free_surface_is_accepted():
    cfg = parse_env(channel_env("water"))
    assert not cfg.unsupported and cfg.water_free_surface_enabled == 0
    s = cfg.to_struct()
    assert s.physics_mode == abi.PHYS_WATER and s.fs_enabled == 0
    cfg = parse_env(channel_env("air", LBM_FREE_SURFACE_MODE="height_kinematic", LBM_FREE_SURFACE_DEPTH_M=6))
    assert not cfg.unsupported and cfg.to_struct().fs_enabled == 1 and cfg.to_struct().physics_mode == abi.PHYS_AIR


@pytest.mark.parametrize("physics", ["air", "water"])
def test_static_bed_links_are_reset_in_air_only(physics):
    """A static bed (flat): in air every closed link becomes
    frac f + (1 - frac) w_q rho0 (rest equilibrium, zero wall velocity); in
    water the closed links keep their values (only a moving bed is
    conditioned there)."""
    cfg, o = make_oracle(channel_env(physics), fp64=True)
    shape = (19, cfg.Nz, cfg.Ny, cfg.Nx)
    rng = np.random.default_rng(3)
    f0 = rng.uniform(0.01, 0.2, size=shape)
    o.upload(abi.FIELD_F, f0)
    o.stage(STAGE_BED_POPULATIONS)
    f1 = o.download(abi.FIELD_F)
    lo = o.download(abi.FIELD_LATTICE_OPEN)
    frac = o.download(abi.FIELD_LATTICE_OPEN_FRAC)
    closed = lo == 0
    assert closed.any() and (~closed).any()
    assert np.array_equal(f1[~closed], f0[~closed])
    if physics == "air":
        expected = frac * f0 + (1.0 - frac) * (W * cfg.rho0)[:, None, None, None]
        assert np.allclose(f1[closed], expected[closed], rtol=0, atol=1e-15)
        # a cut link above the bed (0 < frac < 1) mixes both
        assert np.any((frac > 0) & (frac < 1) & closed)
    else:
        assert np.array_equal(f1[closed], f0[closed])
    o.close()


def test_air_with_a_free_surface_tracks_the_column():
    """Air physics with a height-function free surface: the surface and the
    fill follow the column (the free-surface rules apply although the mode
    is air)."""
    cfg, o = make_oracle(channel_env("air", LBM_FREE_SURFACE_MODE="height_static", LBM_FREE_SURFACE_DEPTH_M=6.4,
                                     LBM_G_LB="1e-4"), fp64=True)
    h = o.download(abi.FIELD_FS_H)
    assert np.allclose(h, cfg.free_surface_base_m)
    fill = o.download(abi.FIELD_FS_FILL)
    ty = o.download(abi.FIELD_FS_TYPE)
    # base 2.3 + 6.4 = 8.7 m: cell 8 is 70 % full, cells above are gas
    # the level is a float32 of the configuration
    assert fill[8, 0, 0] == pytest.approx(0.7, abs=1e-6) and ty[8, 0, 0] == abi.FS_INTERFACE
    assert np.all(ty[9:] == abi.FS_GAS) and np.all(fill[3:8] == 1.0)
    o.step(5)
    assert np.all(np.isfinite(o.download(abi.FIELD_F)))
    # static height mode: the surface does not move
    assert np.array_equal(o.download(abi.FIELD_FS_H), h)
    o.close()
