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


def test_det_exp_against_libm():
    import math

    from oracle_ref import load_oracle

    lib = load_oracle()
    xs = np.concatenate([np.linspace(-40, 40, 2001), [0.0, 1.0, -1.0, 1e-9, 700.0, -700.0]])
    for x in xs:
        e = lib.orc_det_exp_f64(float(x))
        assert e == pytest.approx(math.exp(x), rel=4e-16, abs=0.0), x
    assert lib.orc_det_exp_f64(0.0) == 1.0


def test_prescribed_solitary_surface_moves_at_the_celerity():
    env = height_env(LBM_FREE_SURFACE_MODE="prescribed_solitary", LBM_FREE_SURFACE_INITIAL="solitary",
                     LBM_SOLITARY_AMPLITUDE_M=2.0, LBM_SOLITARY_X0_M=8.0, LBM_X_BOUNDARY="periodic")
    cfg, o = make_oracle(env, fp64=True)
    t = 4.0 / cfg.solitary_celerity_phys  # the crest travels 4 m
    o.stage(STAGE_ADVECT, t)
    h = o.download(abi.FIELD_FS_H)[0].astype(np.float64)
    x = np.arange(cfg.Nx) + 0.5
    kappa = np.sqrt(3.0 * 2.0 / (4.0 * 12.0**3))
    xc = 8.0 + cfg.solitary_celerity_phys * np.float32(t)
    xrel = x - xc
    xrel = np.where(xrel > 0.5 * cfg.Nx, xrel - cfg.Nx, xrel)  # folded into the periodic domain
    xrel = np.where(xrel < -0.5 * cfg.Nx, xrel + cfg.Nx, xrel)
    expect = cfg.free_surface_base_m + 2.0 / np.cosh(kappa * xrel) ** 2
    np.testing.assert_allclose(h, expect, rtol=0, atol=1e-5)
    assert abs(x[np.argmax(h)] - 12.0) <= 0.5
    o.close()


def test_free_surface_population_projection():
    """initialize_free_surface_boundary_populations: at rest (u = 0) the
    closure replaces a population streamed from gas by
    feq_q(rho_fs) + feq_qbar(rho_fs) - feq_qbar(rho) = f_q + 2 w_q (rho_fs - rho);
    populations with wet sources are unchanged."""
    env = height_env(LBM_PRESSURE_FORMULATION="total_pressure")
    _, plain = make_oracle(env, fp64=True)
    _, proj = make_oracle(dict(env, LBM_INIT_FREE_SURFACE_POPULATIONS="1"), fp64=True)
    f0 = plain.download(abi.FIELD_F)
    f1 = proj.download(abi.FIELD_F)
    rho = plain.download(abi.FIELD_RHO)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
    top = 11  # the last wet plane (surface at 12 m)
    # links from above (cz < 0: q 6, 13, 14, 17, 18) read gas; rho_fs = rho0 in the total-pressure form
    for q in (6, 13, 14, 17, 18):
        np.testing.assert_allclose(f1[q, top], f0[q, top] + 2.0 * w[q] * (1.0 - rho[top]), atol=1e-12)
    for q in (0, 1, 2, 3, 4, 5):
        assert np.array_equal(f1[q, top], f0[q, top])
    assert np.array_equal(f1[:, :top], f0[:, :top])
    plain.close()
    proj.close()
