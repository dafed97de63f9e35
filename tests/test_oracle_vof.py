"""Pins of the oracle's VOF pipeline (CODEX:1305-2374) against invariants,
closed forms and an independent connectivity labelling."""
import numpy as np
import pytest
from scipy import ndimage

import lbm_abi as abi
from harness import env_of, make_oracle
from lbm_config import parse_env
from oracle_ref import OracleCase


def vof_env(**kw):
    base = dict(LBM_LX_M=32, LBM_LY_M=16, LBM_LZ_M=40, LBM_FREE_SURFACE_DEPTH_M=20, LBM_VOF_BLOCK_SIZE_M=8,
                LBM_VOF_BLOCK_BOTTOM_M=28, LBM_G_LB=1e-3, LBM_VOF_COLLAPSE_INTERVAL=25,
                LBM_VOF_COLLAPSE_MAX_VOLUME_M3=64)
    base.update(kw)
    return env_of("vof_192x96x160", **base)


def test_flat_pool_at_rest_stays_flat():
    """Flat pool, no block: the component-balanced hydrostatic state is a
    fixed point (WATER_FREE_SURFACE_IMPLEMENTATION.md:184-192: flat water
    stays flat with constant volume)."""
    env = vof_env(LBM_VOF_INITIAL_SHAPE="none", LBM_PRESSURE_FORMULATION="vof_component_balanced")
    cfg, o = make_oracle(env, fp64=True)
    t0 = o.download(abi.FIELD_FS_TYPE)
    v0 = o.reduce(abi.REDUCE_VOF_VOLUME)
    o.step(40)
    assert np.array_equal(o.download(abi.FIELD_FS_TYPE), t0)
    assert abs(o.reduce(abi.REDUCE_VOF_VOLUME) - v0) < 1e-9
    assert np.max(np.abs(o.download(abi.FIELD_UZ))) < 1e-10
    h = o.download(abi.FIELD_FS_H)
    assert np.allclose(h, cfg.free_surface_base_m, atol=1e-9)
    o.close()


def test_falling_block_mass_bounded():
    cfg, o = make_oracle(vof_env(), fp64=True)
    m0 = o.reduce(abi.REDUCE_VOF_MASS)
    o.step(150)
    m1 = o.reduce(abi.REDUCE_VOF_MASS)
    assert abs(m1 - m0) / m0 < 1e-4
    ty = o.download(abi.FIELD_FS_TYPE)
    assert set(np.unique(ty)) <= {abi.FS_GAS, abi.FS_LIQUID, abi.FS_INTERFACE}
    # every liquid cell next to gas is an interface cell (enforce_vof_interface_layer)
    liquid = ty == abi.FS_LIQUID
    gas = ty == abi.FS_GAS
    struct = ndimage.generate_binary_structure(3, 2)  # the 18 D3Q19 neighbours
    gas_near = ndimage.binary_dilation(gas, structure=struct)
    assert not np.any(liquid & gas_near)
    o.close()


def _pocket_state(cfg, pocket):
    """Pool (fill 1 below the base level) with a sealed gas pocket."""
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    fill = np.zeros((nz, ny, nx), np.float32)
    fill[: int(cfg.free_surface_base_m)] = 1.0
    (k0, k1), (j0, j1), (i0, i1) = pocket
    fill[k0:k1, j0:j1, i0:i1] = 0.0
    ctype = np.where(fill > 0.5, abi.FS_LIQUID, abi.FS_GAS).astype(np.int32)
    # interface layer at the free surface and around the pocket
    gas = ctype == abi.FS_GAS
    near = ndimage.binary_dilation(gas, structure=ndimage.generate_binary_structure(3, 2)) & ~gas
    ctype[near] = abi.FS_INTERFACE
    fill[near] = 0.9
    return fill, ctype


def test_trapped_gas_pocket_collapse_closed_form():
    """A sealed pocket of V cells is the trapped-gas candidate with volume
    V dx^3 and mass V rho; it collapses to liquid and the same mass leaves
    the exterior interface cells (CODEX:2058-2126)."""
    env = vof_env(LBM_VOF_INITIAL_SHAPE="none", LBM_VOF_COLLAPSE_INTERVAL=1, LBM_VOF_COLLAPSE_MAX_VOLUME_M3=64)
    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct(), fp64=True)
    from harness import setup_case

    setup_case(o, cfg)
    pocket = ((8, 11), (6, 9), (10, 13))  # 27 cells
    fill, ctype = _pocket_state(cfg, pocket)
    o.upload(abi.FIELD_FS_FILL, fill)
    o.upload(abi.FIELD_FS_TYPE, ctype)
    rho = o.download(abi.FIELD_RHO)
    mass = np.where(ctype == abi.FS_GAS, 0.0, fill * rho)
    o.upload(abi.FIELD_FS_MASS, mass)
    o.set_step(1)
    o.stage(28)
    c = o.counters()
    # the pocket (27 cells, gas fraction 1) and its interface shell (the 90
    # D3Q19 neighbours, gas fraction 1 - 0.9) are unconnected to the top
    (k0, k1), (j0, j1), (i0, i1) = pocket
    shell = np.zeros_like(ctype, bool)
    shell[k0 - 1:k1 + 1, j0 - 1:j1 + 1, i0 - 1:i1 + 1] = ctype[k0 - 1:k1 + 1, j0 - 1:j1 + 1, i0 - 1:i1 + 1] == abi.FS_INTERFACE
    assert shell.sum() == 90
    vol = 27.0 + 90 * (1.0 - float(np.float32(0.9)))
    assert c["free_surface_vof_collapse_candidate_volume_m3"] == pytest.approx(vol, rel=1e-12)
    assert c["free_surface_vof_collapse_applied_volume_m3"] == pytest.approx(vol, rel=1e-12)
    assert c["free_surface_vof_collapse_cell_count"] == 27 + 90
    ty = o.download(abi.FIELD_FS_TYPE)
    assert np.all(ty[k0:k1, j0:j1, i0:i1] == abi.FS_LIQUID)
    assert np.all(ty[shell] == abi.FS_LIQUID)
    m_after = o.download(abi.FIELD_FS_MASS)
    # mass moved into the trapped region equals the mass removed from the
    # exterior interface cells
    trapped = shell.copy()
    trapped[k0:k1, j0:j1, i0:i1] = True
    gain = m_after[trapped].sum() - mass[trapped].sum()
    ext = o.download(abi.FIELD_VOF_EXTERIOR_GAS) == 1
    ext_loss = (mass[ext] - m_after[ext]).sum()
    assert gain == pytest.approx(ext_loss, rel=1e-6)
    o.close()


def test_flood_fill_matches_connectivity_labels():
    """Exterior-gas marks = cells with gas fraction > 1e-6 6-connected to
    the gas of the top layer (fixed point of CODEX:2033-2055)."""
    env = vof_env(LBM_VOF_INITIAL_SHAPE="none", LBM_VOF_COLLAPSE_INTERVAL=1, LBM_VOF_COLLAPSE_MAX_VOLUME_M3=0)
    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct(), fp64=True)
    from harness import setup_case

    setup_case(o, cfg)
    rng = np.random.default_rng(4)
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    gasmask = rng.random((nz, ny, nx)) < 0.45
    ctype = np.where(gasmask, abi.FS_GAS, abi.FS_LIQUID).astype(np.int32)
    fill = np.where(gasmask, 0.0, 1.0).astype(np.float64)
    o.upload(abi.FIELD_FS_TYPE, ctype)
    o.upload(abi.FIELD_FS_FILL, fill)
    o.set_step(1)
    o.stage(28)
    marks = o.download(abi.FIELD_VOF_EXTERIOR_GAS) == 1
    # solid x and y walls: no wrap; 6-connected components of the gas mask
    # that reach the top layer
    assert cfg.x_boundary_mode == "solid" and cfg.y_boundary_mode == "solid"
    lab, _ = ndimage.label(gasmask, structure=ndimage.generate_binary_structure(3, 1))
    seeds = np.unique(lab[nz - 1][lab[nz - 1] > 0])
    expected = np.isin(lab, seeds)
    assert 0 < expected.sum() < gasmask.sum()  # both exterior and trapped gas exist
    assert np.array_equal(marks, expected)
    o.close()


def test_redistribution_conserves_mass():
    """redistribute_vof_mass_excess moves signed excess between cells; with
    the retained excess counted the total is unchanged (CODEX:1417-1574)."""
    cfg, o = make_oracle(vof_env(), fp64=True)
    o.step(60)
    o.stage(20)
    o.stage(21)
    m_before = o.download(abi.FIELD_FS_MASS).sum() + o.download(abi.FIELD_FS_MASS_EXCESS).sum()
    o.stage(22)
    m_after = o.download(abi.FIELD_FS_MASS).sum() + o.download(abi.FIELD_FS_MASS_EXCESS).sum()
    assert m_after == pytest.approx(m_before, rel=1e-12)
    o.close()
