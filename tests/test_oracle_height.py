"""Pins of the oracle's height-function free-surface tracking
(classify_free_surface_cells CODEX:1209-1232, compute_free_surface_column_flux
and advect_free_surface_height CODEX:2376-2429): the cell fill of a known
column height, the column flux against a direct sum, conservation of the
total column volume by the flux-form update (periodic and walled x), the
direction of the change, and a pool at rest staying flat."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import make_oracle

STAGE_CLASSIFY, STAGE_FLUX, STAGE_ADVECT = 32, 33, 34


def height_env(**kw):
    env = {
        "LBM_PHYSICS_MODE": "water", "LBM_FREE_SURFACE_MODE": "height_kinematic", "LBM_FREE_SURFACE_TRACKING": "height",
        "LBM_FREE_SURFACE_INITIAL": "flat", "LBM_COLLISION": "regularized", "LBM_USE_LES": "1", "LBM_TAU": "0.55",
        "LBM_DX_M": "1", "LBM_LX_M": "32", "LBM_LY_M": "8", "LBM_LZ_M": "24", "LBM_BED_PROFILE": "flat",
        "LBM_BED_LEVEL_M": "0", "LBM_FREE_SURFACE_DEPTH_M": "12", "LBM_X_BOUNDARY": "periodic", "LBM_G_LB": "1e-4",
        "LBM_PRESSURE_FORMULATION": "hydrostatic_balanced",
    }
    env.update({k: str(v) for k, v in kw.items()})
    return env


def test_classify_fill_of_known_heights():
    cfg, o = make_oracle(height_env(), fp64=True)
    h = np.full((cfg.Ny, cfg.Nx), 12.0)
    h[:, 5] = 12.3       # 30 % of cell 12
    h[:, 6] = 12.00005   # below the 1e-4 threshold: gas
    h[:, 7] = 12.99995   # within 1e-4 of full: liquid
    o.upload(abi.FIELD_FS_H, h)
    o.stage(STAGE_CLASSIFY)
    fill = o.download(abi.FIELD_FS_FILL)
    ty = o.download(abi.FIELD_FS_TYPE)
    assert fill[12, 0, 5] == pytest.approx(0.3, abs=1e-12) and ty[12, 0, 5] == abi.FS_INTERFACE
    assert ty[12, 0, 6] == abi.FS_GAS and ty[12, 0, 7] == abi.FS_LIQUID
    assert np.all(fill[:12, :, 4] == 1.0) and np.all(fill[12:, :, 4] == 0.0)
    obst = o.download(abi.FIELD_OBSTACLE)
    assert np.all(ty[obst != 0] == abi.FS_GAS)  # closed cells are gas
    o.close()


def _set_flow(o, cfg, ux_col):
    ux = np.broadcast_to(ux_col[None, None, :], (cfg.Nz, cfg.Ny, cfg.Nx)).astype(np.float64)
    o.upload(abi.FIELD_UX, ux)
    o.upload(abi.FIELD_UY, np.zeros_like(ux))


@pytest.mark.parametrize("xb", ["periodic", "solid"])
def test_height_update_conserves_volume_and_has_the_right_sign(xb):
    cfg, o = make_oracle(height_env(LBM_X_BOUNDARY=xb), fp64=True)
    x = np.arange(cfg.Nx) + 0.5
    ux_col = 1e-3 * np.sin(2.0 * np.pi * x / cfg.Nx)  # converges at x = L/2, diverges at x = 0 / L
    h = np.full((cfg.Ny, cfg.Nx), 12.5)
    o.upload(abi.FIELD_FS_H, h)
    o.stage(STAGE_CLASSIFY)
    _set_flow(o, cfg, ux_col)
    o.stage(STAGE_FLUX)
    o.stage(STAGE_ADVECT)
    h1 = o.download(abi.FIELD_FS_H)
    # flux form: the total column volume is unchanged
    assert abs(h1.sum() - h.sum()) < 1e-9 * h.sum()
    dh = h1 - h
    # dh/dt = -dq/dx: the surface rises where the flow converges
    assert dh[0, cfg.Nx // 2] > 0 and dh[0, 1] < 0
    # the column flux is the wet-depth-weighted velocity: 12.5 cells
    qx = 12.5 * ux_col * cfg.vel_scale * cfg.dx
    i = cfg.Nx // 4
    if xb == "periodic":
        expect = -cfg.dt_phys * (qx[i + 1] - qx[i - 1]) / (2.0 * cfg.dx)
        assert dh[0, i] == pytest.approx(expect, rel=1e-6)  # dt, vel_scale are float32 in the config
    o.close()


def test_uniform_flux_leaves_height():
    cfg, o = make_oracle(height_env(), fp64=True)
    o.upload(abi.FIELD_FS_H, np.full((cfg.Ny, cfg.Nx), 12.5))
    o.stage(STAGE_CLASSIFY)
    _set_flow(o, cfg, np.full(cfg.Nx, 2e-3))
    o.stage(STAGE_FLUX)
    o.stage(STAGE_ADVECT)
    assert np.allclose(o.download(abi.FIELD_FS_H), 12.5, atol=1e-12)
    o.close()


def test_pool_at_rest_stays_flat():
    """Flat pool at rest, hydrostatic-balanced pressure: a fixed point (the
    run record WATER_FREE_SURFACE_IMPLEMENTATION.md:184-192)."""
    cfg, o = make_oracle(height_env(), fp64=True)
    ty0 = o.download(abi.FIELD_FS_TYPE)
    o.step(40)
    h = o.download(abi.FIELD_FS_H)
    assert np.allclose(h, cfg.free_surface_base_m, atol=1e-9)
    assert np.array_equal(o.download(abi.FIELD_FS_TYPE), ty0)
    assert np.max(np.abs(o.download(abi.FIELD_UZ))) < 1e-9
    o.close()
