"""Parity of the HIP step (liblbm.so, through the C ABI) with the CPU oracle.

Bar: flags, link fractions, wall types, bed kinematics and VOF cell types
bit-exact; FP32 populations / velocity / density within max-relative 1e-5
after 100 steps (the device and the oracle evaluate the same operations in
the same order without contraction, so they are expected to agree to the
bit; the tolerance is the stated bound); FP16/BF16 storage held to the
same bar against the oracle's FP32 build with the same storage round trip,
and FP16-storage velocity within 1e-3 of the FP64 run after 1000 steps.
"""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of, setup_case
from lbm_config import parse_env
from oracle_ref import OracleCase

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5


def device_case(env):
    from lbm_native import Case

    cfg = parse_env(env)
    c = Case(cfg.to_struct())
    setup_case(c, cfg)
    return cfg, c


def oracle_case(env, fp64=False):
    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct(), fp64=fp64)
    setup_case(o, cfg)
    return cfg, o


def relerr(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-30))


PLAIN_SMALL = dict(LBM_LX_M=32, LBM_LY_M=16, LBM_LZ_M=16)
WAVES_SMALL = dict(LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2, LBM_BED_WAVELENGTH_M=32)
CASES = {
    "plain": env_of("plain_64x32x32", **PLAIN_SMALL),
    "plain_kbc_solidx": env_of("plain_64x32x32", LBM_COLLISION="kbc", LBM_X_BOUNDARY="solid", LBM_ENABLE_TOP_SPONGE=1,
                               LBM_TOP_BOUNDARY="keep", **PLAIN_SMALL),
    "waves": env_of("waves_256x128x128", **WAVES_SMALL),
    "waves_reg_periodic": env_of("waves_256x128x128", LBM_COLLISION="regularized", LBM_X_BOUNDARY="periodic",
                                 **WAVES_SMALL),
    "reference_air": {"LBM_LX_M": "64", "LBM_LY_M": "16", "LBM_LZ_M": "24", "LBM_DX_M": "1", "LBM_BACKSCATTER": "0.02",
                      "LBM_SEED": "7", "LBM_TAU": "0.55"},
    "solid_y": env_of("plain_64x32x32", LBM_Y_BOUNDARY="solid", **PLAIN_SMALL),
    # wall velocity of a closed link taken at the cell centre instead of the
    # link endpoint (CODEX:937-988)
    "waves_cell_wall_velocity": env_of("waves_256x128x128", LBM_WALL_VELOCITY_SAMPLING="cell", **WAVES_SMALL),
}
GEOM_FIELDS = [
    abi.FIELD_LATTICE_OPEN,
    abi.FIELD_LATTICE_OPEN_FRAC,
    abi.FIELD_WALL_TYPE,
    abi.FIELD_NEAR_OBSTACLE,
    abi.FIELD_OBSTACLE,
    abi.FIELD_PHI,
    abi.FIELD_WATER_H,
    abi.FIELD_U_WAVE,
    abi.FIELD_V_WAVE,
    abi.FIELD_W_WAVE,
]


@pytest.mark.parametrize("name", ["plain", "waves", "reference_air", "solid_y"])
def test_geometry_bit_exact(name):
    env = CASES[name]
    _, o = oracle_case(env)
    _, c = device_case(dict(env, LBM_KERNEL_PATH="staged"))
    for steps in (0, 7):
        if steps:
            o.step(steps)
            c.step(steps)
        for fid in GEOM_FIELDS:
            a, b = c.download(fid), o.download(fid)
            assert np.array_equal(a, b), (name, steps, fid, int(np.sum(a != b)))
    c.close()
    o.close()


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("path", ["staged", "fused"])
def test_step_parity_with_oracle(name, path):
    env = dict(CASES[name], LBM_KERNEL_PATH=path)
    cfg = parse_env(env)
    _, o = oracle_case(env)
    _, c = device_case(env)
    steps = 100
    o.step(steps)
    c.step(steps)
    report = {}
    for fid, label in [(abi.FIELD_F, "f"), (abi.FIELD_RHO, "rho"), (abi.FIELD_UX, "ux"), (abi.FIELD_UY, "uy"),
                       (abi.FIELD_UZ, "uz"), (abi.FIELD_OMEGA, "omega")]:
        a, b = c.download(fid), o.download(fid)
        assert np.all(np.isfinite(a)), label
        report[label] = (relerr(a, b), int(np.sum(a != b)))
        assert report[label][0] <= F32_TOL, (name, path, label, report[label])
    print(name, path, report)
    assert c.counters()["boundary_cut_link_count"] == o.counters()["boundary_cut_link_count"]
    assert c.reduce(abi.REDUCE_MASS) == pytest.approx(o.reduce(abi.REDUCE_MASS), rel=1e-12)
    c.close()
    o.close()


@pytest.mark.parametrize("name", ["plain", "waves"])
def test_fused_equals_staged_bitwise(name):
    env = CASES[name]
    _, s = device_case(dict(env, LBM_KERNEL_PATH="staged"))
    _, f = device_case(dict(env, LBM_KERNEL_PATH="fused"))
    for chunk in (3, 5):
        s.step(chunk)
        f.step(chunk)
        for fid in (abi.FIELD_F, abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UZ):
            a, b = f.download(fid), s.download(fid)
            assert np.array_equal(a, b), (name, fid, int(np.sum(a != b)), relerr(a, b))
    s.close()
    f.close()


@pytest.mark.parametrize("opts", [{"block_x": 64}, {"block_x": 128}, {"block_x": 256}])
def test_fused_launch_shapes_bitwise(opts):
    """The block shape of the fused passes changes the schedule only."""
    env = dict(CASES["waves"], LBM_KERNEL_PATH="fused")
    _, a = device_case(env)
    _, b = device_case(env)
    for key, value in opts.items():
        b.set_option(key, value)
    a.step(12)
    b.step(12)
    for fid in (abi.FIELD_F, abi.FIELD_UX, abi.FIELD_OMEGA):
        assert np.array_equal(a.download(fid), b.download(fid)), (opts, fid)
    a.close()
    b.close()


def test_fp64_parity_with_fp64_oracle():
    env = dict(CASES["waves"], LBM_STORAGE="fp64")
    _, o = oracle_case(env, fp64=True)
    _, c = device_case(env)
    o.step(60)
    c.step(60)
    for fid in (abi.FIELD_F, abi.FIELD_UX, abi.FIELD_RHO):
        a, b = c.download(fid), o.download(fid)
        assert a.dtype == np.float64
        assert relerr(a, b) <= 1e-12
    c.close()
    o.close()


@pytest.mark.parametrize("name", ["plain", "waves"])
@pytest.mark.parametrize("storage,rounding", [("fp16", "stochastic"), ("fp16", "nearest"), ("bf16", "stochastic")])
def test_reduced_storage_parity_with_oracle(name, storage, rounding):
    """FP16/BF16 storage: the FP32 oracle applies the same storage round trip
    (same shift, scale and dither), so the device run is held to the FP32
    bar."""
    env = dict(CASES[name], LBM_STORAGE=storage, LBM_STORAGE_ROUNDING=rounding)
    _, o = oracle_case(env)
    _, c = device_case(env)
    o.step(100)
    c.step(100)
    for fid in (abi.FIELD_F, abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UZ):
        a, b = c.download(fid), o.download(fid)
        assert relerr(a, b) <= F32_TOL, (name, storage, fid, int(np.sum(a != b)), relerr(a, b))
    c.close()
    o.close()


@pytest.mark.parametrize("storage,tol", [("fp16", 1e-3), ("bf16", 1e-2)])
def test_reduced_storage_velocity_error(storage, tol):
    """FP16 / BF16 storage (stochastic rounding, mean-flow shift) against the
    FP64 run of the plain channel 64x32x32 after 1000 steps:
    max |u - u64| / max |u64| <= tol, mass drift <= 1e-5."""
    base = env_of("plain_64x32x32")
    _, ref = device_case(dict(base, LBM_STORAGE="fp64"))
    _, c = device_case(dict(base, LBM_STORAGE=storage))
    ref.step(1000)
    c.step(1000)
    u = np.stack([ref.download(x) for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)])
    du = np.stack([c.download(x) for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)]) - u
    vel_err = float(np.max(np.linalg.norm(du, axis=0)) / np.max(np.linalg.norm(u, axis=0)))
    dp = float(np.max(np.abs(c.download(abi.FIELD_RHO) - ref.download(abi.FIELD_RHO))) / 3.0)
    m64 = ref.reduce(abi.REDUCE_MASS)
    dm = abs(c.reduce(abi.REDUCE_MASS) - m64) / m64
    print(storage, "velocity rel error", vel_err, "max |dp|", dp, "mass drift", dm)
    assert vel_err <= tol
    assert dm <= 1e-5
    ref.close()
    c.close()


@pytest.mark.parametrize("name", ["plain", "waves"])
def test_fp32_drift_over_1000_steps(name):
    """FP32 fused path against the FP32 oracle over 1000 steps: the drift
    (max-relative difference of populations, velocity and density) is
    reported at 100, 500 and 1000 steps and held to the 100-step bar; the
    two evaluate the same operations in the same order, so it is 0."""
    env = dict(CASES[name], LBM_KERNEL_PATH="fused")
    _, o = oracle_case(env)
    _, c = device_case(env)
    done = 0
    drift = {}
    for mark in (100, 500, 1000):
        o.step(mark - done)
        c.step(mark - done)
        done = mark
        drift[mark] = {label: relerr(c.download(fid), o.download(fid))
                       for fid, label in [(abi.FIELD_F, "f"), (abi.FIELD_UX, "ux"), (abi.FIELD_RHO, "rho")]}
    print(name, "FP32 drift vs oracle", drift)
    for mark, d in drift.items():
        assert max(d.values()) <= F32_TOL, (name, mark, d)
    c.close()
    o.close()


@pytest.mark.parametrize("storage,bound", [("fp32", 1e-5), ("fp16", 1e-5), ("bf16", 1e-5)])
def test_mass_conservation_closed_channel(storage, bound):
    """Total fluid mass (FP64 device reduction) of the closed plain channel
    (x, y periodic, bed at 1.25 cells with cut links, free-slip top: no mass
    enters or leaves) after 1000 steps: |m - m0| / m0 <= bound.  The FP64
    run of the same case shows what the interpolated bounce-back itself
    changes (cut links with fractions 0.25 / 0.75 are not exactly
    conservative): <= 1e-8."""
    env = env_of("plain_64x32x32", LBM_STORAGE=storage)
    _, c = device_case(env)
    m0 = c.reduce(abi.REDUCE_MASS)
    c.step(1000)
    m1 = c.reduce(abi.REDUCE_MASS)
    drift = abs(m1 - m0) / m0
    print(storage, "mass drift over 1000 steps", drift)
    assert drift <= bound
    c.close()
    if storage == "fp32":
        _, d = device_case(dict(env, LBM_STORAGE="fp64"))
        m0 = d.reduce(abi.REDUCE_MASS)
        d.step(1000)
        d64 = abs(d.reduce(abi.REDUCE_MASS) - m0) / m0
        print("fp64 mass drift over 1000 steps", d64)
        assert d64 <= 1e-8
        d.close()


def test_reference_state_round_trip():
    """Download/upload of the post-stream state continues identically."""
    env = CASES["waves"]
    _, a = device_case(dict(env, LBM_KERNEL_PATH="fused"))
    _, b = device_case(dict(env, LBM_KERNEL_PATH="fused"))
    a.step(10)
    b.step(10)
    b.upload(abi.FIELD_F, b.download(abi.FIELD_F))
    a.step(10)
    b.step(10)
    assert np.array_equal(a.download(abi.FIELD_F), b.download(abi.FIELD_F))
    a.close()
    b.close()


def test_bad_arguments_fail_loudly():
    from lbm_native import LbmError

    _, c = device_case(CASES["plain"])
    small = np.zeros(3, np.float32)
    with pytest.raises(LbmError, match="expected"):
        c._check(c.lib.lbm_upload_field(c.h, abi.FIELD_RHO, small.ctypes.data, small.nbytes), "upload")
    with pytest.raises(LbmError):
        c._check(c.lib.lbm_step(c.h, -1), "step")
    c.close()
