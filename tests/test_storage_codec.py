"""Pins of the reduced-precision population storage (FP16 / BF16).

The oracle's FP32 build applies the device's storage round trip
f -> s + round((f - s) 2^k) 2^-k wherever the device stores populations.
Here the round trip (through an upload of F) is checked against independent
conversions: numpy's IEEE binary16 cast and torch's bfloat16 cast for
round-to-nearest-even, and the two-neighbour / mean property for stochastic
rounding.  The accuracy of the reduced forms against the FP64 run is checked
on the plain channel (config of the FP16 precision study, reduced height)."""
from pathlib import Path

import numpy as np
import pytest
import torch

import lbm_abi as abi
from harness import env_of, setup_case
from lbm_config import parse_env
from oracle_ref import OracleCase, round16

W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float64).astype(np.float32)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0], np.float32)
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, -1, 1], np.float32)
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1], np.float32)


def shifts(u_ref):
    """s_q = feq_q(1, u_ref) in float32, the equilibrium's operation order."""
    f = np.float32
    ux, uy, uz = (f(x) for x in u_ref)
    u2 = ux * ux + uy * uy + uz * uz
    cu = f(3.0) * (CX * ux + CY * uy + CZ * uz)
    return (W * (f(1.0) + cu + f(0.5) * cu * cu - f(1.5) * u2)).astype(np.float32)


def oracle_for(storage, rounding, **kw):
    env = env_of("plain_64x32x32", LBM_LX_M=32, LBM_LY_M=16, LBM_LZ_M=8, LBM_STORAGE=storage,
                 LBM_STORAGE_ROUNDING=rounding, **kw)
    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct())
    setup_case(o, cfg)
    return cfg, o


def probe_populations(cfg, s, scale, seed=3):
    """Populations whose scaled deviations span 2^-30 .. 2^13 (half's
    subnormal range through its top binades), both signs."""
    rng = np.random.default_rng(seed)
    shape = (19, cfg.Nz, cfg.Ny, cfg.Nx)
    mag = np.exp2(rng.uniform(-30.0, 13.0, size=shape))
    d = (mag * rng.choice([-1.0, 1.0], size=shape)).astype(np.float32)
    return (s[:, None, None, None] + d / np.float32(scale)).astype(np.float32)


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_nearest_round_trip_matches_library_casts(storage):
    cfg, o = oracle_for(storage, "nearest")
    s = shifts(cfg.storage_u_ref)
    scale = np.float32(2.0**cfg.storage_shift)
    inv = np.float32(1.0) / scale
    f = probe_populations(cfg, s, scale)
    o.upload(abi.FIELD_F, f)
    got = o.download(abi.FIELD_F)
    d = ((f - s[:, None, None, None]) * scale).astype(np.float32)
    if storage == "fp16":
        h = d.astype(np.float16).astype(np.float32)
    else:
        h = torch.from_numpy(d).to(torch.bfloat16).to(torch.float32).numpy()
    want = (s[:, None, None, None] + h * inv).astype(np.float32)
    assert np.array_equal(got, want)
    o.close()


def neighbours_fp16(d):
    """The two binary16 values around each float32 d: the one toward zero
    and the next one away from zero (equal when d is a half value)."""
    h = d.astype(np.float16)
    zero = np.float16(0)
    toward0 = np.where(np.abs(h.astype(np.float32)) > np.abs(d), np.nextafter(h, zero), h).astype(np.float16)
    outward = np.where(d < 0, np.float16(-np.inf), np.float16(np.inf)).astype(np.float16)
    away = np.where(toward0.astype(np.float32) == d, toward0, np.nextafter(toward0, outward)).astype(np.float16)
    return toward0.astype(np.float32), away.astype(np.float32)


def test_stochastic_rounding_picks_a_neighbour_and_is_unbiased():
    cfg, o = oracle_for("fp16", "stochastic")
    s = shifts(cfg.storage_u_ref)
    scale = np.float32(2.0**cfg.storage_shift)
    inv = np.float32(1.0) / scale
    # one scaled deviation per population, 30 % of an ulp above a half value
    shape = (19, cfg.Nz, cfg.Ny, cfg.Nx)
    base = np.float16(1.5) * np.float16(2.0) ** np.arange(-6, 13, dtype=np.float16)[:19]
    ulp = np.spacing(base).astype(np.float32)
    dq = (base.astype(np.float32) + np.float32(0.3) * ulp) * np.where(np.arange(19) % 2, -1, 1).astype(np.float32)
    d = np.broadcast_to(dq[:, None, None, None], shape).astype(np.float32)
    f = (s[:, None, None, None] + d * inv).astype(np.float32)
    d_exact = ((f - s[:, None, None, None]) * scale).astype(np.float32)
    o.upload(abi.FIELD_F, f)
    got = o.download(abi.FIELD_F)
    lo, hi = neighbours_fp16(d_exact)
    want_lo = (s[:, None, None, None] + lo * inv).astype(np.float32)
    want_hi = (s[:, None, None, None] + hi * inv).astype(np.float32)
    up = got == want_hi
    assert np.all(up | (got == want_lo))
    n = up[0].size
    with np.errstate(invalid="ignore"):
        frac = (np.abs(d_exact) - np.abs(lo)) / (np.abs(hi) - np.abs(lo))
    checked = 0
    for q in range(19):
        if lo[q].flat[0] == hi[q].flat[0]:
            continue  # the float32 population already lies on a half value
        p = float(frac[q].flat[0])
        assert abs(up[q].mean() - p) < 5.0 * np.sqrt(p * (1 - p) / n), q
        checked += 1
    assert checked >= 12
    o.close()


def test_zero_reference_velocity_gives_weights_shift():
    """u_ref = 0: s_q = w_q exactly (the plain f_i - w_i form)."""
    cfg, o = oracle_for("fp16", "nearest", LBM_STORAGE_U_REF="0")
    assert cfg.storage_u_ref == (0.0, 0.0, 0.0)
    assert np.array_equal(shifts(cfg.storage_u_ref), W)
    f = np.broadcast_to(W[:, None, None, None], (19, cfg.Nz, cfg.Ny, cfg.Nx)).astype(np.float32)
    o.upload(abi.FIELD_F, f)
    assert np.array_equal(o.download(abi.FIELD_F), f)
    o.close()


def velocity_error(a, b):
    ua = np.stack([a.download(x) for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)]).astype(np.float64)
    ub = np.stack([b.download(x) for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)]).astype(np.float64)
    return float(np.max(np.linalg.norm(ua - ub, axis=0)) / np.max(np.linalg.norm(ub, axis=0)))


@pytest.mark.parametrize("storage,rounding,tol", [("fp16", "stochastic", 1e-3), ("bf16", "stochastic", 1e-2)])
def test_reduced_storage_error_vs_fp64(storage, rounding, tol):
    """Plain channel 64x32x16, 200 steps: max |u - u64| / max |u64|."""
    env = env_of("plain_64x32x32", LBM_LZ_M=16, LBM_STORAGE=storage, LBM_STORAGE_ROUNDING=rounding)
    cfg = parse_env(env)
    ref = OracleCase(cfg.to_struct(), fp64=True)
    red = OracleCase(cfg.to_struct())
    setup_case(ref, cfg)
    setup_case(red, cfg)
    ref.step(200)
    red.step(200)
    err = velocity_error(red, ref)
    assert err <= tol, err
    m64 = ref.reduce(abi.REDUCE_MASS)
    assert abs(red.reduce(abi.REDUCE_MASS) - m64) / m64 <= 1e-5
    ref.close()
    red.close()


GOLDEN_SR = Path(__file__).resolve().parent / "golden" / "sr_hw_vectors.npz"


def test_stochastic_rounding_matches_hardware_vectors():
    """The oracle's stochastic roundings reproduce, bit for bit, gfx950's
    v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32 (which the device encode issues) on
    the recorded vectors: binary16 normal and subnormal range, underflow,
    overflow next to 65504, the float32 range, zeros, infinities and NaN
    (recorded by tests/golden/make_sr_vectors.py)."""
    v = np.load(GOLDEN_SR)
    x, r = v["x"], v["r"]
    nan = np.isnan(x)
    h = round16(True, True, x, r)
    hb = h.astype(np.float16).view(np.uint16)
    assert np.array_equal(hb[~nan], v["f16"][~nan])
    assert np.isnan(v["f16"][nan].view(np.float16)).all() and np.isnan(h[nan]).all()
    b = round16(False, True, x, r)
    bb = (b.view(np.uint32) >> 16).astype(np.uint16)
    assert np.array_equal(bb[~nan], v["bf16"][~nan])
    # the subnormal range is rounded at the binary16 ulp 2^-24, not at the
    # float's 13th bit: some vectors there differ from add-then-round-nearest
    sub = (np.abs(x) < 2.0**-14) & (np.abs(x) > 0)
    u = x.view(np.uint32)
    naive = (((u & 0x7FFFFFFF) + (r >> 19)) & ~np.uint32(0x1FFF)) | (u & 0x80000000)
    naive16 = naive.view(np.float32).astype(np.float16).view(np.uint16)
    assert (naive16[sub] != v["f16"][sub]).any()


def _on_fp16_grid(f, sh, k):
    """True where a population value v is a stored value: v = fl(s_q + h 2^-k)
    for the binary16 h nearest to (v - s_q) 2^k (float32 arithmetic)."""
    scale = np.float32(2.0**k)
    d = (f - sh[:, None, None, None]) * scale
    h = d.astype(np.float16).astype(np.float32)
    return (sh[:, None, None, None] + h * np.float32(2.0**-k)).astype(np.float32) == f


def test_staged_fp16_stores_the_post_stream_populations():
    """On the staged path (here: a VOF run) a 16-bit run stores the
    post-stream populations, so every population the oracle holds after a
    step lies on the storage grid.  The fused path stores only the
    post-collision set: streaming moves stored values unchanged, but the
    populations it blends (bounce-back at cut links, the inflow and outflow
    layers) leave the grid."""
    from test_oracle_vof import vof_env

    cfg = parse_env(vof_env(LBM_STORAGE="fp16"))
    o = OracleCase(cfg.to_struct())
    setup_case(o, cfg)
    o.step(3)
    f = o.download(abi.FIELD_F)
    assert np.all(_on_fp16_grid(f, shifts(cfg.storage_u_ref), cfg.storage_shift))
    o.close()
    cfg = parse_env(env_of("waves_256x128x128", LBM_LX_M=32, LBM_LY_M=8, LBM_LZ_M=16, LBM_BED_AMP_M=2,
                           LBM_BED_WAVELENGTH_M=16, LBM_STORAGE="fp16"))
    o = OracleCase(cfg.to_struct())
    setup_case(o, cfg)
    o.step(3)
    f = o.download(abi.FIELD_F)
    assert not np.all(_on_fp16_grid(f, shifts(cfg.storage_u_ref), cfg.storage_shift))
    o.close()
