"""The library's host init (lbm_initial_fields, csrc/lbm_init.cpp, the host
half of CODEX:3777-3792) against the Python host init (lbm_init.py), on the
CPU: every array bit for bit, over the initial states (the reference's
random kick and the two new noisy states, whose NumPy default_rng draws the
C++ PCG64 + SeedSequence must reproduce for one- and two-word seeds), the
free-surface initial conditions, the linear-wave pressure on both x
boundaries, the VOF block, the solitary velocity modes, FP64 storage and a
piecewise bed.  The Python side is pinned by tests/test_init_host.py."""
import ctypes

import numpy as np
import pytest

import lbm_abi as abi
from lbm_config import parse_env, preset_env
from lbm_init import initial_fields
from lbm_native import LIB_PATH, initial_fields_native, load_library
from oracle_ref import OracleCase

pytestmark = pytest.mark.skipif(not LIB_PATH.exists(), reason="liblbm.so not built")


def water(**kw):
    env = dict(preset_env("vof_192x96x160"), LBM_LX_M="48", LBM_LY_M="4", LBM_LZ_M="24", LBM_BED_LEVEL_M="2",
               LBM_FREE_SURFACE_DEPTH_M="12", LBM_VOF_INITIAL_SHAPE="none", LBM_PRESSURE_FORMULATION="total_pressure",
               LBM_VOF_BLOCK_SIZE_M="6", LBM_VOF_BLOCK_BOTTOM_M="16", LBM_VOF_COLLAPSE_INTERVAL="25",
               LBM_VOF_COLLAPSE_MAX_VOLUME_M3="64")
    env.update({k: str(v) for k, v in kw.items()})
    return env


def air(**kw):
    env = dict(preset_env("plain_64x32x32"), LBM_LX_M="32", LBM_LY_M="16", LBM_LZ_M="24")
    env.update({k: str(v) for k, v in kw.items()})
    return env


CASES = {
    "uniform_noise": air(),
    "uniform_noise_seed0": air(LBM_SEED=0),
    "uniform_noise_seed_2words": air(LBM_SEED=2**40 + 7),
    "reference_kick": air(LBM_INITIAL_STATE="reference"),
    "reference_kick_fp64": air(LBM_INITIAL_STATE="reference", LBM_STORAGE="fp64", LBM_SEED=99),
    "loglaw_inflow": air(LBM_INITIAL_STATE="loglaw_noise", LBM_X_BOUNDARY="inflow_outflow"),
    "moving_bed": dict(preset_env("waves_256x128x128"), LBM_LX_M="64", LBM_LY_M="16", LBM_LZ_M="32"),
    "water_solitary_divfree": water(LBM_FREE_SURFACE_INITIAL="solitary", LBM_SOLITARY_AMPLITUDE_M=2.0,
                                    LBM_SOLITARY_X0_M=16.0, LBM_X_BOUNDARY="solid"),
    "water_solitary_shallow_periodic": water(LBM_FREE_SURFACE_INITIAL="solitary", LBM_SOLITARY_AMPLITUDE_M=1.0,
                                             LBM_SOLITARY_X0_M=40.0, LBM_X_BOUNDARY="periodic",
                                             LBM_SOLITARY_INITIAL_VELOCITY="shallow_water"),
    "water_solitary_none_fp64": water(LBM_FREE_SURFACE_INITIAL="solitary", LBM_SOLITARY_INITIAL_VELOCITY="none",
                                      LBM_STORAGE="fp64"),
    "water_flat": water(LBM_FREE_SURFACE_INITIAL="flat"),
    "water_gauss_linear_wave_solid": water(LBM_X_BOUNDARY="solid", LBM_FREE_SURFACE_INITIAL="gaussian",
                                           LBM_INITIAL_PRESSURE_MODE="linear_wave", LBM_GAUSSIAN_AMPLITUDE_M=1.0,
                                           LBM_GAUSSIAN_SIGMA_M=5.0),
    "water_gauss_linear_wave_periodic_odd": water(LBM_X_BOUNDARY="periodic", LBM_LX_M="47",
                                                  LBM_FREE_SURFACE_INITIAL="gaussian",
                                                  LBM_INITIAL_PRESSURE_MODE="linear_wave",
                                                  LBM_PRESSURE_FORMULATION="hydrostatic_balanced",
                                                  LBM_GAUSSIAN_AMPLITUDE_M=1.0, LBM_GAUSSIAN_SIGMA_M=5.0),
    "water_gauss_hydrostatic_component": water(LBM_FREE_SURFACE_INITIAL="gaussian", LBM_GAUSSIAN_AMPLITUDE_M=1.5,
                                               LBM_PRESSURE_FORMULATION="vof_component_balanced"),
    "water_vof_block": water(LBM_VOF_INITIAL_SHAPE="block", LBM_GRAVITY_PHYS="0"),
    "water_piecewise_bed": water(LBM_BED_PROFILE="piecewise_linear", LBM_FREE_SURFACE_INITIAL="solitary",
                                 LBM_SOLITARY_X0_M=20.0, LBM_SOLITARY_INITIAL_VELOCITY="shallow_water_divfree"),
}
# bed ramp 2 m -> 6 m between x = 16 m and x = 40 m (written per test)
BOTTOM = "0 2\n16 2\n40 6\n48 6\n"


def host_state(env):
    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct(), fp64=cfg.storage == "fp64")
    if cfg.bed_profile == "piecewise_linear":
        o.upload(abi.FIELD_PIECEWISE_BED_Z, cfg.piecewise_bed_z_np)
    o.geometry(0.0)
    return cfg, o.download(abi.FIELD_OBSTACLE), o.download(abi.FIELD_WATER_H)


@pytest.mark.parametrize("name", list(CASES))
def test_library_init_matches_python_bit_for_bit(name, tmp_path):
    env = dict(CASES[name])
    if env.get("LBM_BED_PROFILE") == "piecewise_linear":
        (tmp_path / "bottom.txt").write_text(BOTTOM)
        env["LBM_BOTTOM_FILE"] = str(tmp_path / "bottom.txt")
    cfg, obstacle0, bed = host_state(env)
    if env.get("LBM_BED_PROFILE") == "piecewise_linear":
        assert np.ptp(bed) > 3.0
    want = initial_fields(cfg, obstacle0, bed)
    got = initial_fields_native(cfg.to_struct(), cfg.init_params(), obstacle0, bed)
    assert set(got) == set(want)
    for fid in want:
        a, b = np.asarray(want[fid]), got[fid]
        assert a.shape == b.shape and a.dtype == b.dtype, fid
        assert np.array_equal(a, b, equal_nan=True), (name, fid, int(np.sum(a != b)))
    # the draws are not constant: the random fields really come from the generator
    if cfg.initial_state != "reference" or not cfg.water_free_surface_enabled:
        ux = got[abi.FIELD_UX]
        assert np.unique(ux).size > ux.size // 2 or cfg.water_free_surface_enabled


def test_flat_default_bed_when_bed_is_null():
    cfg, obstacle0, _ = host_state(CASES["water_flat"])
    want = initial_fields(cfg, obstacle0, None)
    got = initial_fields_native(cfg.to_struct(), cfg.init_params(), obstacle0, None)
    for fid in want:
        assert np.array_equal(np.asarray(want[fid]), got[fid], equal_nan=True), fid


def test_errors():
    lib = load_library()
    cfg = parse_env(CASES["uniform_noise"])
    c, p = cfg.to_struct(), cfg.init_params()
    n = cfg.Nx * cfg.Ny * cfg.Nz
    obst = np.zeros(n, np.int32)
    bufs = [np.empty(n, np.float32) for _ in range(4)] + [np.empty(cfg.Nz, np.float32),
                                                          np.empty(cfg.Nz * cfg.Ny, np.float32)]
    ptrs = [b.ctypes.data for b in bufs]

    def call(params, *ps):
        return lib.lbm_initial_fields(ctypes.byref(c), ctypes.byref(params), obst.ctypes.data, None, *ps, None, None, None)

    assert call(p, *ptrs) == abi.LBM_OK
    assert call(p, None, *ptrs[1:]) == abi.LBM_E_ARG
    bad = abi.LbmInitParams.from_buffer_copy(p)
    bad.version = abi.LBM_INIT_PARAMS_VERSION + 1
    assert call(bad, *ptrs) == abi.LBM_E_CONFIG
    # a water case needs the fill, type and surface buffers
    wc = parse_env(CASES["water_flat"])
    cw, pw = wc.to_struct(), wc.init_params()
    nw = wc.Nx * wc.Ny * wc.Nz
    ob = np.zeros(nw, np.int32)
    wb = [np.empty(nw, np.float32) for _ in range(4)] + [np.empty(wc.Nz, np.float32), np.empty(wc.Nz * wc.Ny, np.float32)]
    rc = lib.lbm_initial_fields(ctypes.byref(cw), ctypes.byref(pw), ob.ctypes.data, None, *[b.ctypes.data for b in wb],
                                None, None, None)
    assert rc == abi.LBM_E_ARG


def test_seed_out_of_range_is_refused():
    cfg = parse_env(air(LBM_SEED=-1))
    with pytest.raises(ValueError, match="LBM_SEED"):
        cfg.init_params()
