"""Device-vs-oracle parity on the device paths that the other GPU tests do
not reach, and on the benchmark configurations at their stated sizes.

Each case asserts that the path under test really ran (its counter, on
both sides).  Bar as in test_gpu_parity / test_gpu_vof: cell types, flags
and labels bit-exact; FP32 fields within max-relative 1e-5 (air) and 1e-6
(VOF) of the FP32 oracle.
  * detached-cell advection and residual compaction (CODEX:1862-1993):
    two isolated droplets far above the pool, one moving, one residual;
  * the thin-gap bridge of the free-surface stream (CODEX:3510-3538): a
    one-cell gas slit through the pool;
  * the refill modes noneq and directional (CODEX:2164-2316);
  * the piecewise-linear bed (CODEX:223-290, 1140-1141) on both paths, with
    a profile whose maximum is far above the bed level;
  * C1 exactly as stated (64x32x32, 500 steps, f / u / dp every 100);
  * C2 at 256x128x128 for 5 steps, C4 at 192x96x160 for 3 steps with a
    sealed gas pocket collapsing on every step.
"""
import numpy as np
import pytest
from scipy import ndimage

import lbm_abi as abi
from harness import env_of, setup_case
from lbm_config import parse_env, preset_env
from oracle_ref import OracleCase
from test_oracle_vof import vof_env

pytestmark = pytest.mark.gpu

GAS, LIQ, INT = abi.FS_GAS, abi.FS_LIQUID, abi.FS_INTERFACE
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, -1, 1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
AIR_FIELDS = [(abi.FIELD_F, "f"), (abi.FIELD_UX, "ux"), (abi.FIELD_UY, "uy"), (abi.FIELD_UZ, "uz"), (abi.FIELD_RHO, "rho"),
              (abi.FIELD_OMEGA, "omega")]
VOF_FIELDS = [abi.FIELD_FS_FILL, abi.FIELD_FS_MASS, abi.FIELD_F, abi.FIELD_UZ]


def relerr(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


def pair(env, prepare=None):
    from lbm_native import Case

    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct())
    c = Case(cfg.to_struct())
    for case in (o, c):
        setup_case(case, cfg)
        if prepare is not None:
            prepare(case, cfg)
    return cfg, o, c


def feq(rho, u):
    cu = CX * u[0] + CY * u[1] + CZ * u[2]
    return rho * W * (1 + 3 * cu + 4.5 * cu**2 - 1.5 * np.dot(u, u))


def with_layer(fill):
    """Cell types of a fill field: liquid / gas, and the interface layer
    (fill 0.9) on the wet side of every wet-gas contact (18 neighbours)."""
    ty = np.where(fill > 0.5, LIQ, GAS).astype(np.int32)
    gas = ty == GAS
    near = ndimage.binary_dilation(gas, structure=ndimage.generate_binary_structure(3, 2)) & ~gas
    ty[near] = INT
    fill = fill.copy()
    fill[near] = 0.9
    return fill, ty


def put_state(case, fill, ty, extra=None):
    rho = case.download(abi.FIELD_RHO)
    if extra is not None:
        extra(case)
    case.upload(abi.FIELD_FS_FILL, fill.astype(np.float32))
    case.upload(abi.FIELD_FS_TYPE, ty)
    case.upload(abi.FIELD_FS_TYPE_PREV, ty)
    case.upload(abi.FIELD_FS_MASS, np.where(ty == GAS, 0.0, fill * rho).astype(np.float32))


def pool_fill(cfg):
    fill = np.zeros((cfg.Nz, cfg.Ny, cfg.Nx), np.float32)
    fill[: int(cfg.free_surface_base_m)] = 1.0
    return fill


def vof_compare(o, c, what):
    assert np.array_equal(c.download(abi.FIELD_FS_TYPE), o.download(abi.FIELD_FS_TYPE)), what
    for fid in VOF_FIELDS:
        assert relerr(c.download(fid), o.download(fid)) <= 1e-6, (what, fid)


def run_vof(env, prepare, steps, keys):
    cfg, o, c = pair(env, prepare)
    seen = {k: 0 for k in keys}
    for s in range(steps):
        o.step(1)
        c.step(1)
        co, cc = o.counters(), c.counters()
        for k in keys:
            assert cc[k] == co[k], (s, k, cc[k], co[k])
            seen[k] += co[k]
        vof_compare(o, c, s)
    o.close()
    c.close()
    return seen


# ---- detached advection / residual compaction --------------------------------
def _droplets(case, cfg):
    fill, ty = with_layer(pool_fill(cfg))

    def pops(cs):
        f = cs.download(abi.FIELD_F)
        for (k, j, i), u in (((34, 8, 8), (0.05, 0.0, 0.08)), ((34, 8, 24), (0.0, 0.03, -0.09))):
            f[:, k, j, i] = feq(1.0, np.array(u))
        cs.upload(abi.FIELD_F, f)

    for (k, j, i), fv in (((34, 8, 8), 0.4), ((34, 8, 24), 0.005)):
        ty[k, j, i] = INT
        fill[k, j, i] = fv
    put_state(case, fill, ty, pops)


def test_detached_advection_and_compaction_parity():
    print("case detached droplets", flush=True)
    seen = run_vof(vof_env(LBM_VOF_INITIAL_SHAPE="none"), _droplets, 3,
                   ["free_surface_vof_detached_advect_count", "free_surface_vof_detached_compress_count",
                    "free_surface_vof_orphan_count", "free_surface_refill_count"])
    assert seen["free_surface_vof_detached_advect_count"] > 0
    assert seen["free_surface_vof_detached_compress_count"] > 0


# ---- thin-gap bridge -----------------------------------------------------------
def _slit(case, cfg):
    fill = pool_fill(cfg)
    fill[4:14, :, 16] = 0.0
    put_state(case, *with_layer(fill))


def test_thin_gap_bridge_parity():
    print("case thin-gap slit", flush=True)
    seen = run_vof(vof_env(LBM_VOF_INITIAL_SHAPE="none", LBM_VOF_THIN_GAP_BRIDGE=1), _slit, 3,
                   ["free_surface_thin_gap_bridge_count", "free_surface_cut_link_count"])
    assert seen["free_surface_thin_gap_bridge_count"] > 0


# ---- refill modes ----------------------------------------------------------------
@pytest.mark.parametrize("mode", ["noneq", "directional"])
def test_refill_modes_parity(mode):
    print("case refill", mode, flush=True)
    cfg, o, c = pair(vof_env(LBM_FREE_SURFACE_REFILL=mode))
    refills = 0
    for chunk in range(6):
        o.step(10)
        c.step(10)
        co, cc = o.counters(), c.counters()
        assert cc["free_surface_refill_count"] == co["free_surface_refill_count"]
        refills += co["free_surface_refill_count"]
        vof_compare(o, c, (mode, chunk))
    assert refills > 0
    o.close()
    c.close()


# ---- piecewise-linear bed ---------------------------------------------------------
@pytest.mark.parametrize("path", ["staged", "fused"])
def test_piecewise_bed_parity(path, tmp_path):
    """A ramp from 1.5 m up to 9.5 m over the domain while the bed level
    (offset) stays 1.5 m: the band of cut links must reach the profile's
    maximum on the fused path."""
    print("case piecewise", path, flush=True)
    bottom = tmp_path / "bottom.txt"
    bottom.write_text("0 1.5\n20 1.5\n48 9.5\n64 9.5\n")
    env = env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_PROFILE="piecewise_linear",
                 LBM_BOTTOM_FILE=str(bottom), LBM_BED_LEVEL_M=1.5, LBM_KERNEL_PATH=path)
    cfg, o, c = pair(env)
    assert cfg.offset_m == 1.5 and float(np.max(cfg.piecewise_bed_z_np)) > 9
    for steps in (0, 40):
        if steps:
            o.step(steps)
            c.step(steps)
        for fid in (abi.FIELD_LATTICE_OPEN, abi.FIELD_LATTICE_OPEN_FRAC, abi.FIELD_WALL_TYPE, abi.FIELD_NEAR_OBSTACLE,
                    abi.FIELD_WATER_H):
            assert np.array_equal(c.download(fid), o.download(fid)), (steps, fid)
    for fid, label in AIR_FIELDS:
        assert relerr(c.download(fid), o.download(fid)) <= 1e-5, (path, label)
    assert c.counters()["boundary_cut_link_count"] == o.counters()["boundary_cut_link_count"] > 0
    o.close()
    c.close()


# ---- C1 as stated -------------------------------------------------------------
def test_c1_as_stated_both_paths():
    """Plain 64x32x32, 500 steps; f, u and dp = (rho - 1)/3 compared every
    100 steps (max-relative 1e-5) on the fused and the staged path."""
    from lbm_native import Case

    print("case C1 64x32x32 x 500", flush=True)
    env = preset_env("plain_64x32x32")
    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct())
    setup_case(o, cfg)
    devs = {}
    for path in ("fused", "staged"):
        e = dict(env, LBM_KERNEL_PATH=path)
        ce = parse_env(e)
        devs[path] = Case(ce.to_struct())
        setup_case(devs[path], ce)
    for mark in range(1, 6):
        o.step(100)
        ref = {label: o.download(fid) for fid, label in AIR_FIELDS[:5]}
        dp_ref = (ref["rho"].astype(np.float64) - 1.0) / 3.0
        for path, c in devs.items():
            c.step(100)
            for fid, label in AIR_FIELDS[:4]:
                assert relerr(c.download(fid), ref[label]) <= 1e-5, (path, 100 * mark, label)
            dp = (c.download(abi.FIELD_RHO).astype(np.float64) - 1.0) / 3.0
            assert float(np.max(np.abs(dp - dp_ref))) <= 1e-5 * max(float(np.max(np.abs(dp_ref))), 1e-30), (path, mark)
    for c in devs.values():
        c.close()
    o.close()


# ---- full-size C2 and C4 ----------------------------------------------------------
def test_c2_full_size_five_steps():
    print("case C2 256x128x128 x 5", flush=True)
    cfg, o, c = pair(preset_env("waves_256x128x128"))
    o.step(5)
    c.step(5)
    for fid, label in AIR_FIELDS:
        assert relerr(c.download(fid), o.download(fid)) <= 1e-5, label
    assert c.counters()["boundary_cut_link_count"] == o.counters()["boundary_cut_link_count"] > 0
    o.close()
    c.close()


def _pocket(case, cfg):
    """The C4 state (pool + falling block) with a sealed 4^3 gas pocket
    carved into the pool."""
    fill = case.download(abi.FIELD_FS_FILL).copy()
    wet = fill > 0.5
    wet[40:44, 40:44, 90:94] = False
    fill = np.where(wet, 1.0, 0.0).astype(np.float32)
    put_state(case, *with_layer(fill))


def test_c4_full_size_with_a_collapsing_pocket():
    print("case C4 192x96x160 x 3, collapse every step", flush=True)
    env = dict(preset_env("vof_192x96x160"), LBM_VOF_COLLAPSE_INTERVAL="1")
    cfg, o, c = pair(env, _pocket)
    collapsed = 0
    for s in range(3):
        o.step(1)
        c.step(1)
        co, cc = o.counters(), c.counters()
        for k in ("free_surface_vof_collapse_cell_count", "free_surface_refill_count", "free_surface_vof_orphan_count"):
            assert cc[k] == co[k], (s, k)
        assert cc["free_surface_vof_collapse_candidate_volume_m3"] == pytest.approx(
            co["free_surface_vof_collapse_candidate_volume_m3"], rel=1e-9)
        collapsed += co["free_surface_vof_collapse_cell_count"]
        assert np.array_equal(c.download(abi.FIELD_VOF_EXTERIOR_GAS), o.download(abi.FIELD_VOF_EXTERIOR_GAS))
        vof_compare(o, c, s)
    assert collapsed > 0
    assert c.reduce(abi.REDUCE_VOF_MASS) == pytest.approx(o.reduce(abi.REDUCE_VOF_MASS), rel=1e-9)
    o.close()
    c.close()


# ---- moving-bed time levels computed several steps at a time ---------------------
@pytest.mark.parametrize("profile", ["moving_sine", "moving_sine1"])
@pytest.mark.parametrize("path", ["fused", "staged"])
def test_bed_levels_across_batches(profile, path):
    """The moving bed's levels are computed up to 8 steps per pair of
    launches into a ring of 9 slots: calls of 1, 8, 9 and 5 steps cross the
    batch and ring boundaries; the bed, its velocities and the populations
    stay equal to the oracle's after each call."""
    print("case bed batches", profile, path, flush=True)
    env = env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2, LBM_BED_WAVELENGTH_M=32,
                 LBM_BED_PROFILE=profile, LBM_KERNEL_PATH=path)
    cfg, o, c = pair(env)
    for n in (1, 8, 9, 5):
        o.step(n)
        c.step(n)
        for fid in (abi.FIELD_WATER_H, abi.FIELD_U_WAVE, abi.FIELD_W_WAVE, abi.FIELD_LATTICE_OPEN_FRAC):
            assert np.array_equal(c.download(fid), o.download(fid)), (n, fid)
        assert relerr(c.download(abi.FIELD_F), o.download(abi.FIELD_F)) <= 1e-5, n
    o.close()
    c.close()
