"""Pins of the oracle's box obstacle (cube_sdf_at_point CODEX:897-934 and
its use in compute_phi_slope_corrected / build_lattice_open, CODEX:2432-2686):
the signed distance against an independent box distance in double, the
yaw by 90 degrees equal to swapped sizes, the link flags and wall types it
produces, and a fluid at rest around the box staying at rest."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of, make_oracle


def cube_env(**kw):
    base = dict(LBM_LX_M=32, LBM_LY_M=16, LBM_LZ_M=24, LBM_OBSTACLE_MODE="cube", LBM_CUBE_SIZE_X_M=8,
                LBM_CUBE_SIZE_Y_M=6, LBM_CUBE_SIZE_Z_M=5, LBM_CUBE_CENTER_X_M=14.3, LBM_CUBE_CENTER_Y_M=8.2,
                LBM_CUBE_CENTER_Z_M=11.6)
    base.update(kw)
    return env_of("plain_64x32x32", **base)


def box_distance(cfg, half, center):
    """Signed distance to an axis-aligned box, double precision, periodic
    in x and y (minimum image)."""
    z, y, x = np.meshgrid(np.arange(cfg.Nz) + 0.5, np.arange(cfg.Ny) + 0.5, np.arange(cfg.Nx) + 0.5, indexing="ij")
    lx, ly = cfg.Nx * cfg.dx, cfg.Ny * cfg.dx
    d = [(x * cfg.dx - center[0] + 0.5 * lx) % lx - 0.5 * lx, (y * cfg.dx - center[1] + 0.5 * ly) % ly - 0.5 * ly,
         z * cfg.dx - center[2]]
    q = [np.abs(d[a]) - half[a] for a in range(3)]
    outside = np.sqrt(sum(np.maximum(v, 0.0) ** 2 for v in q))
    inside = np.minimum(np.maximum(q[0], np.maximum(q[1], q[2])), 0.0)
    return outside + inside


def test_box_distance_closed_form():
    cfg, o = make_oracle(cube_env(), fp64=True)
    phi = o.download(abi.FIELD_PHI)
    expect = box_distance(cfg, (4.0, 3.0, 2.5), (14.3, 8.2, 11.6))
    np.testing.assert_allclose(phi, expect, atol=1e-5)
    o.close()


def test_yaw_quarter_turn_swaps_sizes():
    _, a = make_oracle(cube_env(LBM_CUBE_YAW_DEG=90), fp64=True)
    _, b = make_oracle(cube_env(LBM_CUBE_SIZE_X_M=6, LBM_CUBE_SIZE_Y_M=8), fp64=True)
    np.testing.assert_allclose(a.download(abi.FIELD_PHI), b.download(abi.FIELD_PHI), atol=1e-5)
    assert np.array_equal(a.download(abi.FIELD_OBSTACLE), b.download(abi.FIELD_OBSTACLE))
    a.close()
    b.close()


def test_flags_and_wall_types():
    cfg, o = make_oracle(cube_env(LBM_OBSTACLE_MODE="bed_cube"))
    lo, wt, lf = (np.moveaxis(o.download(f), 0, -1)
                  for f in (abi.FIELD_LATTICE_OPEN, abi.FIELD_WALL_TYPE, abi.FIELD_LATTICE_OPEN_FRAC))
    d = box_distance(cfg, (4.0, 3.0, 2.5), (14.3, 8.2, 11.6))
    inside = d < -1.0  # every link endpoint (half a cell away) is inside
    assert np.all(lo[..., 0][inside] == 0) and np.all(lf[..., 1:][inside] == 0)
    far = (d > 1.5) & (np.arange(cfg.Nz)[:, None, None] > 4)
    assert np.all(lo[far] == 1) and np.all(lf[far][:, 1:] == 1.0)
    # links cut by the box carry its wall type; links cut by the bed the bed's
    cut = (lo[..., 1:] == 0) & (lo[..., :1] == 1)
    kk = np.arange(cfg.Nz)[:, None, None, None]
    near_box = np.abs(d)[..., None] < 1.0
    assert np.all(wt[..., 1:][cut & near_box] == abi.WALL_CUBE)
    assert np.all(wt[..., 1:][cut & (kk < 3)] == abi.WALL_BED)
    o.close()


def test_rest_state_around_the_box_stays_at_rest():
    env = cube_env(LBM_INIT_U_LU=0, LBM_INIT_NOISE_LU=0, LBM_CUBE_YAW_DEG=30, LBM_CUBE_PITCH_DEG=20)
    cfg, o = make_oracle(env, fp64=True)
    o.step(30)
    u = np.stack([o.download(f) for f in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)])
    assert np.max(np.abs(u)) < 1e-12
    o.close()


def test_phi_modes_on_a_flat_bed_equal_the_vertical_gap():
    """On a flat bed phi is the vertical distance z - h, and its trilinear
    sample at a link endpoint is exact for a linear field: the phi modes
    reproduce the vertical mode's fractions and flags."""
    base = env_of("plain_64x32x32", LBM_LX_M=16, LBM_LY_M=8, LBM_LZ_M=12, LBM_BED_LEVEL_M=3.3)
    _, v = make_oracle(base, fp64=True)
    for mode in ("phi_fraction", "phi"):
        _, p = make_oracle(dict(base, LBM_BOUNDARY_GEOMETRY=mode), fp64=True)
        np.testing.assert_allclose(p.download(abi.FIELD_LATTICE_OPEN_FRAC), v.download(abi.FIELD_LATTICE_OPEN_FRAC),
                                   atol=1e-12)
        assert np.array_equal(p.download(abi.FIELD_LATTICE_OPEN), v.download(abi.FIELD_LATTICE_OPEN))
        p.close()
    v.close()


def test_phi_fraction_on_a_slope_is_the_normal_distance(tmp_path):
    """On a uniform slope s the phi field is the normal distance
    (z - h) / sqrt(1 + s^2), so away from the clamps every open fraction of
    the phi modes is 0.5 + (frac_vertical - 0.5) / sqrt(1 + s^2)."""
    nx = 32
    slope = 0.25
    bottom = tmp_path / "bottom.txt"
    bottom.write_text(f"0 2.0\n{nx} {2.0 + slope * nx}\n")
    env = env_of("plain_64x32x32", LBM_LX_M=nx, LBM_LY_M=4, LBM_LZ_M=16, LBM_BED_PROFILE="piecewise_linear",
                 LBM_BOTTOM_FILE=str(bottom), LBM_X_BOUNDARY="inflow_outflow")
    _, v = make_oracle(env, fp64=True)
    _, p = make_oracle(dict(env, LBM_BOUNDARY_GEOMETRY="phi_fraction"), fp64=True)
    lf_v = v.download(abi.FIELD_LATTICE_OPEN_FRAC)[5]  # +z link: its endpoint lies above the cell centre
    lf_p = p.download(abi.FIELD_LATTICE_OPEN_FRAC)[5]
    inner = (lf_v > 0.05) & (lf_v < 0.95)
    inner[:, :, :2] = False
    inner[:, :, -2:] = False
    assert inner.sum() > 10
    expect = 0.5 + (lf_v[inner] - 0.5) / np.sqrt(1.0 + slope**2)
    np.testing.assert_allclose(lf_p[inner], expect, atol=2e-6)
    v.close()
    p.close()
