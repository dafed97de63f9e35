"""Parity of the device VOF pipeline (staged path) with the oracle.

Bar: cell types and exterior-gas labels bit-exact; populations, fill and
mass within max-relative 1e-6 (the gather formulation of the scatters is
shared, and every operation matches the oracle's order); the FP64
collapse sums within 1e-9 relative (parallel vs serial summation)."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of, setup_case
from lbm_config import parse_env
from oracle_ref import OracleCase
from test_oracle_vof import _pocket_state, vof_env

pytestmark = pytest.mark.gpu


def pair(env):
    from lbm_native import Case

    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct())
    setup_case(o, cfg)
    c = Case(cfg.to_struct())
    setup_case(c, cfg)
    return cfg, o, c


def relerr(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def test_vof_initial_state_matches():
    cfg, o, c = pair(vof_env())
    for fid in (abi.FIELD_FS_TYPE, abi.FIELD_FS_FILL, abi.FIELD_FS_MASS, abi.FIELD_F, abi.FIELD_FS_H):
        assert np.array_equal(c.download(fid), o.download(fid)), fid
    c.close()
    o.close()


def test_vof_falling_block_parity():
    cfg, o, c = pair(vof_env())
    for chunk in (40, 40, 40, 30):
        o.step(chunk)
        c.step(chunk)
        assert np.array_equal(c.download(abi.FIELD_FS_TYPE), o.download(abi.FIELD_FS_TYPE))
        for fid in (abi.FIELD_FS_FILL, abi.FIELD_FS_MASS, abi.FIELD_F, abi.FIELD_UZ):
            assert relerr(c.download(fid), o.download(fid)) <= 1e-6, (c.step_index(), fid)
    cc, oc = c.counters(), o.counters()
    for key in ("free_surface_refill_count", "free_surface_vof_orphan_count", "free_surface_cut_link_count",
                "boundary_cut_link_count", "free_surface_coalescence_repair_count"):
        assert cc[key] == oc[key], key
    m_dev, m_ref = c.reduce(abi.REDUCE_VOF_MASS), o.reduce(abi.REDUCE_VOF_MASS)
    assert m_dev == pytest.approx(m_ref, rel=1e-9)
    c.close()
    o.close()


def test_vof_trapped_gas_collapse_parity():
    env = vof_env(LBM_VOF_INITIAL_SHAPE="none", LBM_VOF_COLLAPSE_INTERVAL=1, LBM_VOF_COLLAPSE_MAX_VOLUME_M3=64)
    cfg, o, c = pair(env)
    pocket = ((8, 11), (6, 9), (10, 13))
    fill, ctype = _pocket_state(cfg, pocket)
    rho = o.download(abi.FIELD_RHO)
    mass = np.where(ctype == abi.FS_GAS, 0.0, fill * rho).astype(np.float32)
    for case in (o, c):
        case.upload(abi.FIELD_FS_FILL, fill)
        case.upload(abi.FIELD_FS_TYPE, ctype)
        case.upload(abi.FIELD_FS_MASS, mass)
    o.step(1)
    c.step(1)
    cc, oc = c.counters(), o.counters()
    assert cc["free_surface_vof_collapse_cell_count"] == oc["free_surface_vof_collapse_cell_count"] > 0
    assert cc["free_surface_vof_collapse_candidate_volume_m3"] == pytest.approx(
        oc["free_surface_vof_collapse_candidate_volume_m3"], rel=1e-9)
    assert cc["flood_fill_unconverged"] == 0
    assert np.array_equal(c.download(abi.FIELD_VOF_EXTERIOR_GAS), o.download(abi.FIELD_VOF_EXTERIOR_GAS))
    assert np.array_equal(c.download(abi.FIELD_FS_TYPE), o.download(abi.FIELD_FS_TYPE))
    assert relerr(c.download(abi.FIELD_FS_MASS), o.download(abi.FIELD_FS_MASS)) <= 1e-6
    c.close()
    o.close()
