"""Pins of the oracle's bed kinematics and boundary rules against closed
forms and invariants (FP64 oracle, one reference kernel at a time through
the oracle's stage entry).

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign, a wrong index range, an offset of half a cell) fails it:
  * the reference's default bed, two harmonics with a smoothstep ramp
    (CODEX:1082-1162), on both sides of the ramp;
  * the x fringe and the top drive of apply_open_boundary_conditions
    (CODEX:3582-3662) on a known post-stream state;
  * the top sponge of collide_KBC (CODEX:3191-3200);
  * the backscatter injection and its gamma limiter (CODEX:3245-3315);
  * the x solid-wall bounce-back (CODEX:3554-3580);
  * the link-endpoint wall velocity of the moving-wall correction
    (sample_wall_velocity_lb, CODEX:937-988, used at CODEX:3674-3712).
Oracle stage numbers: 0 bed, 5 macro, 7 collision, 8 stream, 9 x walls,
10 open boundaries, 11 swap f / f_new.
"""
import math

import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of
from lbm_config import parse_env
from oracle_ref import OracleCase

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, -1, 1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = [0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12, 17, 18, 15, 16]
EPS = 1e-6


def f32(x):
    """A config value as the device and the oracle see it (stored as float)."""
    return float(np.float32(x))


def feq(rho, ux, uy, uz):
    """D3Q19 equilibrium (19, ...) in FP64 from its moments definition:
    rho w (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u^2)."""
    rho, ux, uy, uz = (np.asarray(a, float) for a in (rho, ux, uy, uz))
    sh = (19,) + np.broadcast(rho, ux, uy, uz).shape
    e = (slice(None),) + (None,) * (len(sh) - 1)
    cu = CX[e] * ux + CY[e] * uy + CZ[e] * uz
    return np.broadcast_to(rho * W[e] * (1 + 3 * cu + 4.5 * cu**2 - 1.5 * (ux**2 + uy**2 + uz**2)), sh)


def oracle(env):
    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct(), fp64=True)
    o.geometry(0.0)
    return cfg, o


def plain(**kw):
    base = dict(LBM_LX_M=16, LBM_LY_M=8, LBM_LZ_M=12, LBM_BED_LEVEL_M=-100)
    base.update(kw)
    return env_of("plain_64x32x32", **base)


def smoothstep(x):
    return x * x * (3 - 2 * x)


# ---- a3: the reference's default bed (CODEX:1082-1162) ----------------------
@pytest.mark.parametrize("frac_of_ramp", [0.3, 0.9, 1.6])
def test_two_harmonic_ramped_bed_closed_form(frac_of_ramp):
    """eta = A(t) [sin(k/2 (x - c t)) + sin(k (x - c t / 2))] with A(t) =
    A smoothstep(t / 2T) during the first two wave periods T and A after;
    u_s = c eta1/(h0 + eta1) + c/2 eta2/(h0 + eta2); w_s = eta_t + u_s eta_x
    with eta_t the backward difference of the stored heights; velocities in
    lattice units."""
    env = {"LBM_LX_M": "32", "LBM_LY_M": "4", "LBM_LZ_M": "16", "LBM_DX_M": "1", "LBM_BED_PROFILE": "moving_sine",
           "LBM_BED_AMP_M": "1.5", "LBM_BED_WAVELENGTH_M": "16"}
    cfg, o = oracle(env)
    s = cfg.to_struct()
    assert s.bed_profile == abi.BED_MOVING_SINE
    A, lam, c, h0 = f32(s.bed_amp_m), f32(s.bed_wavelength_m), f32(s.wave_speed), f32(s.depth_swl_m)
    ramp = 2 * f32(s.wave_period)
    dt = f32(s.dt_phys)
    t = float(np.float32(frac_of_ramp * ramp))
    t_prev = float(np.float32(t - dt))
    o.stage(0, t_prev)
    o.stage(0, t)
    k = 2 * math.pi / lam
    x = np.arange(cfg.Nx) * f32(s.dx)

    def bed(tt):
        a_now = A * smoothstep(tt / ramp) if tt < ramp else A
        e1 = a_now * np.sin(0.5 * k * (x - c * tt))
        e2 = a_now * np.sin(k * (x - 0.5 * c * tt))
        ex = a_now * 0.5 * k * np.cos(0.5 * k * (x - c * tt)) + a_now * k * np.cos(k * (x - 0.5 * c * tt))
        return e1, e2, ex

    e1, e2, ex = bed(t)
    p1, p2, _ = bed(t_prev)
    off = f32(s.offset_m)
    h = o.download(abi.FIELD_WATER_H)
    assert np.allclose(h, (off + e1 + e2)[None, :], rtol=0, atol=1e-12)
    us = c * e1 / (h0 + e1) + 0.5 * c * e2 / (h0 + e2)
    ws = ((off + e1 + e2) - (off + p1 + p2)) / dt + us * ex
    vs = f32(s.vel_scale)
    assert np.allclose(o.download(abi.FIELD_U_WAVE), (us / vs)[None, :], rtol=1e-10, atol=1e-14)
    assert np.allclose(o.download(abi.FIELD_W_WAVE), (ws / vs)[None, :], rtol=1e-7, atol=1e-10)
    assert np.all(o.download(abi.FIELD_V_WAVE) == 0)
    # inside the ramp the full-amplitude bed is not what the oracle holds
    full = A * (np.sin(0.5 * k * (x - c * t)) + np.sin(k * (x - 0.5 * c * t)))
    assert np.allclose(h, (off + full)[None, :], rtol=0, atol=1e-12) == (t >= ramp)
    o.close()


# ---- a15: x fringe and top drive (CODEX:3582-3662) --------------------------
def test_x_fringe_and_top_drive_closed_form():
    """On a post-stream state A: the cells i < 5 of the upper half
    (k > Nz/2) relax to feq(rho0, u_inlet[k], 0, 0) with sigma = 0.05
    ((k / (Nz-1))^2 - 1/4) smoothstep(1 - i/4); every other cell is left as
    it is; then the top plane is overwritten by feq(rho0, u_top, 0, 0)."""
    env = plain(LBM_ENABLE_X_FRINGE=1, LBM_ENABLE_TOP_DRIVE=1, LBM_LZ_M=14)
    cfg, o = oracle(env)
    s = cfg.to_struct()
    assert s.enable_x_fringe == 1 and s.enable_top_drive == 1
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    rng = np.random.default_rng(21)
    A = W[:, None, None, None] * (1 + 0.05 * rng.uniform(-1, 1, (19, nz, ny, nx)))
    u_in = 0.04 + 0.002 * np.arange(nz)
    o.upload(abi.FIELD_U_INLET_PROFILE, u_in)
    o.upload(abi.FIELD_F, A)
    o.stage(11)  # f_new <- A
    o.upload(abi.FIELD_F, A)
    o.stage(10)
    got = o.download_fnew()
    rho0 = f32(s.rho0)
    exp = A.copy()
    for k in range(nz // 2 + 1, nz):
        for i in range(5):
            sig = 0.05 * ((k / (nz - 1)) ** 2 - 0.25) * smoothstep(1 - i / 4)
            exp[:, k, :, i] = (1 - sig) * A[:, k, :, i] + sig * feq(rho0, u_in[k], 0.0, 0.0)[:, None]
    exp[:, nz - 1] = feq(rho0, f32(s.u_top_lb), 0.0, 0.0)[:, None, None]
    assert np.allclose(got, exp, rtol=0, atol=1e-15)
    # the fringe really acted (and only where it should)
    assert not np.allclose(got[:, nz - 2, :, 0], A[:, nz - 2, :, 0])
    assert np.array_equal(got[:, : nz // 2 + 1], A[:, : nz // 2 + 1])
    assert np.array_equal(got[:, : nz - 1, :, 5:], A[:, : nz - 1, :, 5:])
    o.close()


# ---- a12: top sponge of collide_KBC (CODEX:3191-3200) -----------------------
def test_top_sponge_is_bgk_at_tau_095():
    """Planes k > Nz - 3 relax with tau = 0.95 and no limiter; the planes
    below take the KBC/BGK step with omega = 1/tau0."""
    env = plain(LBM_COLLISION="kbc", LBM_ENABLE_TOP_SPONGE=1, LBM_USE_LES=0, LBM_TOP_BOUNDARY="keep")
    cfg, o = oracle(env)
    s = cfg.to_struct()
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    rng = np.random.default_rng(4)
    f = W[:, None, None, None] * (1 + 0.02 * rng.standard_normal((19, nz, ny, nx)))
    o.upload(abi.FIELD_F, f)
    o.stage(5)
    rho = o.download(abi.FIELD_RHO)
    u = [o.download(x) for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)]
    o.stage(7)
    fp = o.download(abi.FIELD_F)
    fe = feq(rho, *u)
    om = 1 / f32(s.tau0)
    sp = slice(nz - 2, nz)
    assert np.allclose(fp[:, sp], f[:, sp] - (1 / 0.95) * (f[:, sp] - fe[:, sp]), atol=1e-15)
    assert np.allclose(fp[:, : nz - 2], f[:, : nz - 2] - om * (f[:, : nz - 2] - fe[:, : nz - 2]), atol=1e-15)
    o.close()


# ---- a12: backscatter (CODEX:3245-3315) -------------------------------------
def _backscatter(c_bs, tau=0.52, seed=11):
    env = plain(LBM_COLLISION="kbc", LBM_BACKSCATTER=c_bs, LBM_USE_LES=0, LBM_TAU=tau, LBM_SEED=seed)
    cfg, o = oracle(env)
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    rng = np.random.default_rng(8)
    f = W[:, None, None, None] * (1 + 0.05 * rng.standard_normal((19, nz, ny, nx)))
    o.upload(abi.FIELD_F, f)
    o.stage(5)
    rho = o.download(abi.FIELD_RHO)
    u = [o.download(x) for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)]
    o.stage(7)
    fp = o.download(abi.FIELD_F)
    o.close()
    fe = feq(rho, *u)
    om = 1 / f32(cfg.to_struct().tau0)
    fk = f - om * (f - fe)  # the KBC step without backscatter (limiter inactive here)
    assert fk.min() > EPS
    return fk, fe, fp


def _deviatoric_basis(fk, fe):
    """b_q = w_q 4.5 Q_q : P_dev with P_dev the trace-free part of the
    post-collision non-equilibrium stress."""
    d = fk - fe
    c = [CX, CY, CZ]
    P = np.array([[np.tensordot(c[a] * c[b], d, 1) for b in range(3)] for a in range(3)])
    tr = (P[0, 0] + P[1, 1] + P[2, 2]) / 3
    for a in range(3):
        P[a, a] = P[a, a] - tr
    b = np.zeros_like(d)
    for q in range(19):
        cq = [c[0][q], c[1][q], c[2][q]]
        contr = sum((cq[a] * cq[b_] - (1 / 3 if a == b_ else 0)) * P[a, b_] for a in range(3) for b_ in range(3))
        b[q] = W[q] * 4.5 * contr
    return b


def test_backscatter_is_a_bounded_deviatoric_kick():
    """The increment over KBC conserves mass and momentum, is trace-free,
    is proportional to the deviatoric projection of the cell's own
    non-equilibrium stress, with a factor |C xi| <= C of both signs, and
    doubles with C (gamma = 1 here)."""
    fk, fe, fp = _backscatter(0.02)
    d = fp - fk
    assert np.allclose(d.sum(0), 0, atol=1e-16)
    for ca in (CX, CY, CZ):
        assert np.allclose(np.tensordot(ca, d, 1), 0, atol=1e-16)
    assert np.allclose(sum(np.tensordot(ca * ca, d, 1) for ca in (CX, CY, CZ)), 0, atol=1e-16)
    b = _deviatoric_basis(fk, fe)
    s = (d * b).sum(0) / (b * b).sum(0)
    assert np.allclose(d, s[None] * b, rtol=0, atol=1e-15)
    assert np.all(np.abs(s) <= 0.02 * (1 + 1e-12))
    assert s.min() < -0.005 and s.max() > 0.005  # xi covers [-1, 1]
    fk2, fe2, fp2 = _backscatter(0.04)
    s2 = ((fp2 - fk2) * b).sum(0) / (b * b).sum(0)
    assert np.allclose(s2, 2 * s, rtol=1e-9, atol=1e-15)
    _, _, fp0 = _backscatter(0.0)
    assert np.allclose(fp0, fk, atol=1e-15)


def test_backscatter_gamma_limiter_stops_at_the_floor():
    """With a huge C the kick is cut by gamma: no population ends below
    1e-6; wherever the cut acts (most cells here) the smallest population
    lands on 1e-6 exactly, and the cut kick keeps the deviatoric shape and
    conserves mass and momentum.  Cells where |xi| is small enough keep the
    full kick (gamma clipped to 1)."""
    fk, fe, fp = _backscatter(1000.0)
    assert fp.min() >= EPS - 1e-15
    hit = np.abs(fp.min(axis=0) - EPS) <= 1e-15
    assert hit.mean() > 0.9
    d = fp - fk
    b = _deviatoric_basis(fk, fe)
    s = (d * b).sum(0) / (b * b).sum(0)
    assert np.allclose(d, s[None] * b, rtol=0, atol=1e-14)
    assert np.all(np.abs(s[hit]) < 1000.0)
    assert np.allclose(d.sum(0), 0, atol=1e-14)
    for ca in (CX, CY, CZ):
        assert np.allclose(np.tensordot(ca, d, 1), 0, atol=1e-14)


# ---- a14: x solid-wall bounce-back (CODEX:3554-3580) ------------------------
def test_x_wall_reflects_a_population_into_its_opposite():
    env = plain(LBM_X_BOUNDARY="solid", LBM_TOP_BOUNDARY="keep")
    cfg, o = oracle(env)
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    for q, i in ((2, 0), (9, 0), (13, 0), (1, nx - 1), (7, nx - 1), (14, nx - 1)):
        if CX[q] < 0:
            assert i == 0
        f = np.zeros((19, nz, ny, nx))
        f[q, 5, 3, i] = 1.0
        o.upload(abi.FIELD_F, f)
        o.stage(8)
        o.stage(9)
        fn = o.download_fnew()
        exp = np.zeros_like(fn)
        exp[OPP[q], 5, 3, i] = 1.0
        assert np.array_equal(fn, exp), q
    o.close()


def test_rest_state_with_solid_x_walls_stays_at_rest():
    env = plain(LBM_X_BOUNDARY="solid", LBM_TOP_BOUNDARY="keep", LBM_INIT_U_LU=0, LBM_INIT_NOISE_LU=0,
                LBM_BED_LEVEL_M=1.25)
    from harness import setup_case

    cfg = parse_env(env)
    o = OracleCase(cfg.to_struct(), fp64=True)
    setup_case(o, cfg)
    o.step(60)
    open_ = o.download(abi.FIELD_OBSTACLE) == 0
    for fid in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ):
        assert np.max(np.abs(o.download(fid)[open_])) < 1e-15
    assert np.max(np.abs(o.download(abi.FIELD_RHO)[open_] - 1)) < 1e-14
    o.close()


# ---- a4: link-endpoint wall velocity (CODEX:937-988, 3674-3712) -------------
@pytest.mark.parametrize("sampling", ["link", "cell"])
def test_wall_velocity_is_sampled_at_the_link_endpoint(sampling):
    """A wall velocity linear in x, u = a + b x, w = a' + b' x (cell values
    at the cell centres): with link sampling the moving-wall correction of
    link q uses the value at x_c + c_x dx / 2 exactly (bilinear sampling is
    exact on a linear field), with cell sampling the value at x_c.  With
    f = f_new = 0 the corrected population is
    6 w_q rho0 (c_x u + c_z w)(1 - frac_q)."""
    env = plain(LBM_BED_LEVEL_M=1.25, LBM_WALL_VELOCITY_SAMPLING=sampling)
    cfg, o = oracle(env)
    s = cfg.to_struct()
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    dx = f32(s.dx)
    xc = (np.arange(nx) + 0.5) * dx
    a, b, a2, b2 = 0.003, 0.0011, -0.002, 0.0007
    o.upload(abi.FIELD_U_WAVE, np.broadcast_to(a + b * xc, (ny, nx)))
    o.upload(abi.FIELD_V_WAVE, np.zeros((ny, nx)))
    o.upload(abi.FIELD_W_WAVE, np.broadcast_to(a2 + b2 * xc, (ny, nx)))
    zero = np.zeros((19, nz, ny, nx))
    o.upload(abi.FIELD_F, zero)
    o.stage(11)
    o.upload(abi.FIELD_F, zero)
    o.stage(10)
    fn = o.download_fnew()
    near = o.download(abi.FIELD_NEAR_OBSTACLE) == 1
    lf = o.download(abi.FIELD_LATTICE_OPEN_FRAC)
    wt = o.download(abi.FIELD_WALL_TYPE)
    rho0 = f32(s.rho0)
    ks, js, is_ = np.nonzero(near)
    assert ks.size > 0
    checked = 0
    for k, j, i in zip(ks, js, is_):
        if i == 0 or i == nx - 1:
            continue  # the periodic wrap of the sampling is not linear
        for q in range(1, 19):
            if wt[q, k, j, i] != abi.WALL_BED or lf[q, k, j, i] >= 1:
                continue
            x = xc[i] + (0.5 * CX[q] * dx if sampling == "link" else 0.0)
            uw, ww = a + b * x, a2 + b2 * x
            exp = 6 * W[q] * rho0 * (CX[q] * uw + CZ[q] * ww) * (1 - lf[q, k, j, i])
            assert abs(fn[q, k, j, i] - exp) < 1e-16, (q, k, j, i, fn[q, k, j, i], exp)
            checked += 1
    assert checked > 0
    o.close()
