"""FP16 / BF16 population storage on the staged path (water, VOF and
height tracking, the box obstacle, the phi modes, a forced staged air run)
against the FP32 oracle that applies the same storage round trips.

The staged step stores populations after the collision and after the
stream, and where the conditioning, the refill and the initial free-surface
projection write; the oracle rounds at the same points with the same
shift, scale and dither (CODEX's populations are consumed at CODEX:1306-1350
by the mass exchange, at CODEX:2164-2316 by the refill and at 2319-2373 by
the coalescence repair).  Bar as for FP32: cell types bit-exact, fields
within max-relative 1e-6 (VOF) / 1e-5 (air, height tracking), counters
equal.

Also the error of the FP16 water run against the FP64 run of the same case
(falling block into a pool, 150 steps): water volume, cell types, velocity
and the pressure perturbation, with the bounds written below."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of
from lbm_config import parse_env
from test_gpu_cube import CASES as BOX_CASES
from test_gpu_height import CASES as HEIGHT_CASES
from test_gpu_parity import F32_TOL, device_case, oracle_case
from test_gpu_vof import pair, relerr
from test_oracle_vof import vof_env

pytestmark = pytest.mark.gpu

VOF_FIELDS = (abi.FIELD_FS_FILL, abi.FIELD_FS_MASS, abi.FIELD_F, abi.FIELD_UZ, abi.FIELD_RHO)


@pytest.mark.parametrize("storage,rounding", [("fp16", "stochastic"), ("fp16", "nearest"), ("bf16", "stochastic")])
def test_vof_falling_block_16bit_parity(storage, rounding):
    cfg, o, c = pair(vof_env(LBM_STORAGE=storage, LBM_STORAGE_ROUNDING=rounding))
    assert c.path() == abi.PATH_STAGED
    refills = 0
    for chunk in (40, 40, 40, 30):
        o.step(chunk)
        c.step(chunk)
        assert np.array_equal(c.download(abi.FIELD_FS_TYPE), o.download(abi.FIELD_FS_TYPE)), (storage, c.step_index())
        for fid in VOF_FIELDS:
            assert relerr(c.download(fid), o.download(fid)) <= 1e-6, (storage, c.step_index(), fid)
        cc, oc = c.counters(), o.counters()
        for key in ("free_surface_refill_count", "free_surface_vof_orphan_count", "free_surface_cut_link_count",
                    "free_surface_coalescence_repair_count"):
            assert cc[key] == oc[key], (storage, key)
        refills += oc["free_surface_refill_count"]
    assert c.reduce(abi.REDUCE_VOF_MASS) == pytest.approx(o.reduce(abi.REDUCE_VOF_MASS), rel=1e-9)
    c.close()
    o.close()


@pytest.mark.parametrize("name", ["mound_periodic", "solitary_solid_x", "mound_projected_init"])
def test_height_tracking_16bit_parity(name):
    env = dict(HEIGHT_CASES[name], LBM_STORAGE="fp16")
    cfg, o, c = pair(env)
    o.step(60)
    c.step(60)
    assert np.array_equal(c.download(abi.FIELD_FS_TYPE), o.download(abi.FIELD_FS_TYPE))
    for fid in (abi.FIELD_FS_H, abi.FIELD_FS_FILL, abi.FIELD_F, abi.FIELD_UX, abi.FIELD_RHO):
        assert relerr(c.download(fid), o.download(fid)) <= F32_TOL, (name, fid)
    assert c.counters()["free_surface_refill_count"] == o.counters()["free_surface_refill_count"]
    c.close()
    o.close()


STAGED_AIR = {
    "waves_forced_staged": env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2,
                                  LBM_BED_WAVELENGTH_M=32, LBM_KERNEL_PATH="staged"),
    "bed_cube_waves": BOX_CASES["bed_cube_waves"],
    "phi_waves": BOX_CASES["phi_waves"],
}


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
@pytest.mark.parametrize("name", list(STAGED_AIR))
def test_staged_air_16bit_parity(name, storage):
    env = dict(STAGED_AIR[name], LBM_STORAGE=storage)
    _, o = oracle_case(env)
    _, c = device_case(env)
    assert c.path() == abi.PATH_STAGED
    o.step(100)
    c.step(100)
    for fid in (abi.FIELD_F, abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UZ, abi.FIELD_OMEGA):
        assert relerr(c.download(fid), o.download(fid)) <= F32_TOL, (name, storage, fid)
    assert c.counters()["boundary_cut_link_count"] == o.counters()["boundary_cut_link_count"]
    c.close()
    o.close()


def water_errors(ref, c, p_scale):
    """Errors of run c against the FP64 run ref: water volume
    |V - V64| / V64; the fraction of cells whose type differs; over the cells
    liquid in both runs, the velocity error max |du| / max |u64| and
    rms |du| / rms |u64|, and the error of the pressure perturbation
    p' = (rho - rho0) / 3 in units of p_scale (max and rms)."""
    t64, t = ref.download(abi.FIELD_FS_TYPE), c.download(abi.FIELD_FS_TYPE)
    liq = (t64 == abi.FS_LIQUID) & (t == abi.FS_LIQUID)
    u = np.stack([ref.download(x) for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)])
    du = np.stack([c.download(x) for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)]) - u
    p = (ref.download(abi.FIELD_RHO) - 1.0) / 3.0
    dp = (c.download(abi.FIELD_RHO) - 1.0) / 3.0 - p
    un, dun = np.linalg.norm(u, axis=0)[liq], np.linalg.norm(du, axis=0)[liq]
    pl, dpl = p[liq], dp[liq]
    v64 = ref.reduce(abi.REDUCE_VOF_VOLUME)
    rms = lambda x: float(np.sqrt(np.mean(np.square(x))))
    return {"volume": abs(c.reduce(abi.REDUCE_VOF_VOLUME) - v64) / v64, "type_mismatch": float(np.mean(t != t64)),
            "velocity_max": float(np.max(dun) / np.max(un)), "velocity_rms": rms(dun) / rms(un),
            "pressure_max": float(np.max(np.abs(dpl))) / p_scale, "pressure_rms": rms(dpl) / p_scale}


def test_vof_fp16_error_against_fp64():
    """FP16 storage of the falling-block case against its FP64 run after 150
    steps (the FP32 run beside it for scale).  Bounds: water volume <= 1e-5;
    cells of a different type <= 1e-3; over the cells liquid in both runs,
    velocity rms <= 2e-3 and max <= 1e-2 of the FP64 field's, pressure
    perturbation max <= 1e-3 of the pool's hydrostatic head rho0 g H (the
    component-balanced p' is a small difference from rho0, so its own
    magnitude is no scale: FP32 rounding of rho alone is 10 % of it).
    Interface cells are left out of the field norms: a cell that is gas in
    one run and wet in the other differs by its whole velocity."""
    env = vof_env()
    cfg = parse_env(env)
    p_scale = float(cfg.g_lb) * float(cfg.free_surface_base_m) / float(cfg.dx)
    _, ref = device_case(dict(env, LBM_STORAGE="fp64"))
    runs = {s: device_case(dict(env, LBM_STORAGE=s))[1] for s in ("fp32", "fp16")}
    ref.step(150)
    for c in runs.values():
        c.step(150)
    err = {s: water_errors(ref, c, p_scale) for s, c in runs.items()}
    for s, e in err.items():
        print(s, e)
    e = err["fp16"]
    assert e["volume"] <= 1e-5
    assert e["type_mismatch"] <= 1e-3
    assert e["velocity_rms"] <= 2e-3 and e["velocity_max"] <= 1e-2
    assert e["pressure_max"] <= 1e-3
    ref.close()
    for c in runs.values():
        c.close()
