"""Host initialisation (prepare_initial_density_field CODEX:743-818,
solitary surface CODEX:1166-1182 and initial velocity CODEX:2756-2800),
pinned against closed forms: the linear-wave pressure of a single surface
mode decays as cosh(k z') / cosh(k h) for both mode families, a uniform
elevation gives a uniform head, the solitary surface is A sech^2(kappa x),
and the divergence-free solitary velocity has a vanishing discrete
divergence up to the O(dx^2) truncation."""
import numpy as np
import pytest

from lbm_config import parse_env
from lbm_init import initial_density, linear_wave_head, solitary_velocity, surface_height


def water_env(**kw):
    env = {
        "LBM_PHYSICS_MODE": "water", "LBM_FREE_SURFACE_MODE": "height_kinematic", "LBM_FREE_SURFACE_TRACKING": "vof",
        "LBM_DX_M": "1", "LBM_LX_M": "64", "LBM_LY_M": "4", "LBM_LZ_M": "32", "LBM_BED_PROFILE": "flat",
        "LBM_BED_LEVEL_M": "2", "LBM_FREE_SURFACE_DEPTH_M": "16", "LBM_G_LB": "1e-5",
    }
    env.update({k: str(v) for k, v in kw.items()})
    return parse_env(env)


@pytest.mark.parametrize("xb,mode", [("solid", 3), ("periodic", 2)])
def test_linear_wave_single_mode(xb, mode):
    cfg = water_env(LBM_X_BOUNDARY=xb)
    nx, dx = cfg.Nx, cfg.dx
    x = (np.arange(nx) + 0.5) * dx
    k = mode * np.pi / (nx * dx) if xb == "solid" else 2.0 * np.pi * mode / (nx * dx)
    a = 0.3
    eta = a * np.cos(k * x)
    bed = 2.0
    head = linear_wave_head(cfg, eta, bed)
    depth = cfg.free_surface_base_m - bed
    for kk in range(cfg.Nz):
        zr = min(max((kk + 0.5) * dx - bed, 0.0), depth)
        expect = a * np.cosh(k * zr) / np.cosh(k * depth) * np.cos(k * x)
        np.testing.assert_allclose(head[kk], expect, atol=1e-12)


def test_linear_wave_uniform_elevation_is_hydrostatic():
    cfg = water_env(LBM_X_BOUNDARY="periodic")
    head = linear_wave_head(cfg, np.full(cfg.Nx, 0.25), 2.0)
    np.testing.assert_allclose(head, 0.25, atol=1e-12)


def test_linear_wave_density_follows_head():
    cfg = water_env(LBM_X_BOUNDARY="solid", LBM_FREE_SURFACE_INITIAL="gaussian", LBM_INITIAL_PRESSURE_MODE="linear_wave",
                    LBM_PRESSURE_FORMULATION="hydrostatic_balanced", LBM_GAUSSIAN_AMPLITUDE_M=1.0,
                    LBM_GAUSSIAN_SIGMA_M=6.0)
    fs_h = surface_height(cfg)
    rho = initial_density(cfg, fs_h, np.full((cfg.Ny, cfg.Nx), 2.0, np.float32))
    head = linear_wave_head(cfg, fs_h.astype(np.float64).mean(axis=0) - cfg.free_surface_base_m, 2.0)
    expect = (cfg.rho0 + cfg.g_lb * head / cfg.dx * 3.0).astype(np.float32)
    assert rho.shape == (cfg.Nz, cfg.Ny, cfg.Nx)
    np.testing.assert_allclose(rho[:, 1, :], expect, rtol=0, atol=2e-7)
    # deep below the crest the pressure perturbation is smaller than at the surface
    crest = int(np.argmax(fs_h[0]))
    assert abs(rho[2, 0, crest] - cfg.rho0) < abs(rho[17, 0, crest] - cfg.rho0)
    # a non-flat bed falls back to the hydrostatic head
    bed = np.full((cfg.Ny, cfg.Nx), 2.0, np.float32)
    bed[0, 0] = 2.5
    rho_h = initial_density(cfg, fs_h, bed)
    assert not np.allclose(rho_h, rho)


def test_solitary_surface_closed_form():
    cfg = water_env(LBM_FREE_SURFACE_INITIAL="solitary", LBM_SOLITARY_AMPLITUDE_M=2.0, LBM_SOLITARY_X0_M=20.0,
                    LBM_X_BOUNDARY="solid")
    h = surface_height(cfg)[0].astype(np.float64)
    x = np.arange(cfg.Nx) + 0.5
    kappa = np.sqrt(3.0 * 2.0 / (4.0 * 16.0**3))
    expect = cfg.free_surface_base_m + 2.0 / np.cosh(kappa * (x - 20.0)) ** 2
    np.testing.assert_allclose(h, expect, rtol=0, atol=2e-5)
    assert int(np.argmax(h)) in (19, 20)


def test_solitary_velocity_is_divergence_free():
    cfg = water_env(LBM_FREE_SURFACE_INITIAL="solitary", LBM_SOLITARY_AMPLITUDE_M=1.0, LBM_SOLITARY_X0_M=32.0,
                    LBM_X_BOUNDARY="solid", LBM_LX_M=128)
    fs_h = surface_height(cfg)
    bed = np.full((cfg.Ny, cfg.Nx), 2.0, np.float32)
    ux, uz = solitary_velocity(cfg, fs_h, bed)
    ux = ux.astype(np.float64) - cfg.u_current_lb
    uz = uz.astype(np.float64)
    # shallow-water velocity at the crest: c A / (h0 + A)
    crest = int(np.argmax(fs_h[0]))
    c = cfg.solitary_celerity_phys
    eta_c = float(fs_h[0, crest]) - cfg.free_surface_base_m
    assert ux[10, 0, crest] * cfg.vel_scale == pytest.approx(c * eta_c / (16.0 + eta_c), rel=1e-5)
    # discrete divergence in the wet interior, relative to |du/dx|
    dudx = (ux[:, :, 2:] - ux[:, :, :-2]) / 2.0
    dwdz = (uz[2:, :, :] - uz[:-2, :, :]) / 2.0
    div = dudx[3:15, :, :] + dwdz[2:14, :, 1:-1]
    assert np.max(np.abs(div)) < 0.05 * np.max(np.abs(dudx))
    # mode "none" adds nothing
    cfg2 = water_env(LBM_FREE_SURFACE_INITIAL="solitary", LBM_SOLITARY_INITIAL_VELOCITY="none")
    u2, w2 = solitary_velocity(cfg2, surface_height(cfg2), bed[:, :cfg2.Nx])
    assert np.all(u2 == np.float32(cfg2.u_current_lb)) and np.all(w2 == 0)
