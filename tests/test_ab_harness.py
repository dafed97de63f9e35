"""The metrics harness (ab_harness.py, the counterpart of run_ab_case.py)
on synthetic module states whose metrics follow in closed form: sampled
sines (exact discrete derivatives and exact mean squares over whole
periods), hand-counted neighbourhoods, the cosh decay of one linear-wave
mode, and the column layout of the metrics file (run_ab_case.py:15-36)."""
import csv
from types import SimpleNamespace

import numpy as np
import pytest

import ab_harness as abh
import lbm_abi as abi
from lbm_solver import D3Q19_CX, D3Q19_CY, D3Q19_CZ


class Arr:
    def __init__(self, a):
        self.a = np.asarray(a)

    def to_numpy(self):
        return self.a.copy()


class Cnt:
    def __init__(self, v):
        self.v = v

    def __getitem__(self, key):
        assert key is None
        return self.v


def fake_module(nx=32, ny=4, nz=24, dx=0.5, xb="periodic", fs=False, **extra):
    m = SimpleNamespace(Nx=nx, Ny=ny, Nz=nz, dx=dx, Lx_m=nx * dx, Ly_m=ny * dx, Lz_m=nz * dx, rho0=1.0,
                        vel_scale=20.0, rho_phys=1.2, rho_air_phys=1.2, gravity_phys=9.81, x_boundary_mode=xb,
                        water_free_surface_enabled=int(fs), free_surface_tracking_mode="height",
                        free_surface_base_m=8.0, use_hydrostatic_balanced_pressure=0,
                        use_component_balanced_pressure=0, vof_empty_fill_threshold=0.05,
                        FS_GAS=abi.FS_GAS, FS_LIQUID=abi.FS_LIQUID, FS_INTERFACE=abi.FS_INTERFACE)
    m.x_phys = (np.arange(nx) + 0.5) * dx
    m.y_phys = (np.arange(ny) + 0.5) * dx
    m.z_phys = (np.arange(nz) + 0.5) * dx
    m.Z_phys, m.Y_phys, m.X_phys = np.meshgrid(m.z_phys, m.y_phys, m.x_phys, indexing="ij")
    m.cx, m.cy, m.cz = Arr(D3Q19_CX), Arr(D3Q19_CY), Arr(D3Q19_CZ)
    shape = (nz, ny, nx)
    m.rho = Arr(np.ones(shape))
    m.ux, m.uy, m.uz = Arr(np.zeros(shape)), Arr(np.zeros(shape)), Arr(np.zeros(shape))
    bed = 1.0
    m.water_h = Arr(np.full((ny, nx), bed))
    m.obstacle = Arr((m.Z_phys < bed).astype(np.int32))
    m.phi = Arr(m.Z_phys - bed)
    for name, v in extra.items():
        setattr(m, name, v)
    return m


def test_column_set_matches_the_metrics_file_format():
    assert len(abh.METRIC_FIELDNAMES) == 55 and len(set(abh.METRIC_FIELDNAMES)) == 55
    m = fake_module()
    m.ux.a[5, 1, 3] = 0.01
    row = abh.frame_metrics(m, "x", 10, 0.5)
    assert set(row) == set(abh.METRIC_FIELDNAMES)


def test_high_pass_zero_on_linear_fields_and_two_on_checkerboard():
    k, i = np.meshgrid(np.arange(12.0), np.arange(10.0), indexing="ij")
    lin = 3.0 * k - 2.0 * i + 7.0
    assert np.max(np.abs(abh.high_pass_xz(lin, False))) < 1e-12
    assert np.max(np.abs(abh.high_pass_xz(5.0 * k + 1.0, True))) < 1e-12
    cb = (-1.0) ** (k + i)
    hp = abh.high_pass_xz(cb, True)
    # interior: f - (1/4)(4 * -f) = 2 f
    assert np.allclose(hp[1:-1], 2.0 * cb[1:-1], atol=1e-12)


def test_vorticity_and_pressure_statistics_of_sampled_sines():
    """ux = B sin(2 pi k / Nz), uz = A sin(2 pi i / Nx), rho = rho0 + e cos(2 pi i / Nx):
    the central differences are exact, A sin(th) cos(th i) - B sin(ph) cos(ph k), and
    whole-period mean squares of sampled cosines are 1/2."""
    m = fake_module()
    nz, ny, nx = m.Nz, m.Ny, m.Nx
    th, ph = 2 * np.pi / nx, 2 * np.pi / nz
    kk, _, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    A, B, e = 0.02, 0.01, 1e-3
    m.uz.a = A * np.sin(th * ii)
    m.ux.a = B * np.sin(ph * kk)
    m.rho.a = 1.0 + e * np.cos(th * ii)
    row = abh.frame_metrics(m, "s", 1, 0.0)
    band_k = [k for k in range(nz) if 0.0 < m.z_phys[k] - 1.0 <= 10 * m.dx]
    vort = A * np.sin(th) * np.cos(th * np.arange(nx))[None, :] - B * np.sin(ph) * np.cos(ph * np.array(band_k))[:, None]
    assert row["vort_rms"] == pytest.approx(np.sqrt(np.mean((vort - vort.mean()) ** 2)), rel=1e-10)
    # pressure perturbation rho_phys vs^2 e cos / 3 over whole x periods
    assert row["p_rms"] == pytest.approx(1.2 * 400.0 * e / 3.0 / np.sqrt(2.0), rel=1e-10)
    assert row["p_total_rms"] == row["p_rms"]
    # the x high-pass of cos(th i) is (1 - cos th)/2 cos(th i); z-constant in the band
    assert row["p_hp_rms"] == pytest.approx(1.2 * 400.0 * e / 3.0 * 0.5 * (1 - np.cos(th)) / np.sqrt(2.0), rel=1e-10)
    assert np.isnan(row["p_linear_resid_rms"]) and np.isnan(row["fs_p_rms"])
    assert row["surface_min"] != row["surface_min"] and row["max_speed_fs_type"] == -1


def test_solid_x_walls_use_one_sided_differences():
    m = fake_module(xb="solid")
    ii = np.broadcast_to(np.arange(m.Nx, dtype=np.float64), (m.Nz, m.Ny, m.Nx))
    m.uz.a = 0.003 * ii  # linear in x: one-sided and central differences are both exact
    row = abh.frame_metrics(m, "s", 1, 0.0)
    assert row["vort_rms"] == pytest.approx(0.0, abs=1e-15)
    assert row["vort_hp_rms"] == pytest.approx(0.0, abs=1e-15)
    m.x_boundary_mode = "periodic"  # the wrap now breaks the two end columns
    assert abh.frame_metrics(m, "s", 1, 0.0)["vort_rms"] > 1e-3


def test_max_speed_location_and_neighbourhood():
    m = fake_module(fs=True, xb="solid")
    nz, ny, nx = m.Nz, m.Ny, m.Nx
    k0 = 10
    ty = np.full((nz, ny, nx), abi.FS_LIQUID, dtype=np.int32)
    ty[k0] = abi.FS_INTERFACE
    ty[k0 + 1:] = abi.FS_GAS
    fill = np.where(ty == abi.FS_LIQUID, 1.0, 0.0)
    fill[k0] = 0.4
    m.free_surface_type, m.free_surface_fill = Arr(ty), Arr(fill)
    m.free_surface_h = Arr(np.full((ny, nx), (k0 + 0.4) * m.dx))
    m.ux.a[k0, 2, 7] = 0.05
    row = abh.frame_metrics(m, "n", 1, 0.0)
    assert (row["max_speed_x"], row["max_speed_y"], row["max_speed_z"]) == (7.5 * m.dx, 2.5 * m.dx, (k0 + 0.5) * m.dx)
    assert row["max_speed_phys"] == pytest.approx(0.05 * m.vel_scale)
    assert row["max_speed_fs_type"] == abi.FS_INTERFACE and row["max_speed_fs_fill"] == 0.4
    # 5 links up (gas), 8 in-plane (fill 0.4), 5 down (fill 1)
    assert (row["max_speed_wet_neighbors"], row["max_speed_full_neighbors"], row["max_speed_gas_neighbors"]) == (13, 5, 5)
    # at the x = 0 wall the 5 links with cx = -1 leave the domain: one up, one down, three in-plane
    m.ux.a[k0, 2, 7] = 0.0
    m.ux.a[k0, 2, 0] = 0.05
    row = abh.frame_metrics(m, "n", 1, 0.0)
    assert (row["max_speed_wet_neighbors"], row["max_speed_full_neighbors"], row["max_speed_gas_neighbors"]) == (9, 4, 4)
    # gas cells are inactive: a faster gas cell is ignored
    m.ux.a[k0 + 3, 1, 1] = 0.5
    assert abh.frame_metrics(m, "n", 1, 0.0)["max_speed_x"] == 0.5 * m.dx


def test_surface_statistics_of_a_sine_surface():
    m = fake_module(fs=True)
    nz, ny, nx = m.Nz, m.Ny, m.Nx
    a, th = 0.3, 2 * np.pi / nx
    eta = a * np.sin(th * (np.arange(nx) + 0.5))
    h = np.broadcast_to(m.free_surface_base_m + eta, (ny, nx)).copy()
    m.free_surface_h = Arr(h)
    z = m.Z_phys
    fill = np.clip((h[None] - (z - 0.5 * m.dx)) / m.dx, 0.0, 1.0)
    ty = np.where(fill >= 1.0, abi.FS_LIQUID, np.where(fill > 0.0, abi.FS_INTERFACE, abi.FS_GAS))
    obs = m.obstacle.a.astype(bool)
    ty[obs] = abi.FS_GAS
    m.free_surface_type, m.free_surface_fill = Arr(ty), Arr(fill)
    m.ux.a[:] = 0.001
    row = abh.frame_metrics(m, "w", 1, 0.0)
    assert row["surface_min"] == pytest.approx(h.min()) and row["surface_max"] == pytest.approx(h.max())
    # the mirror image of a sine over the period is its negative: l2 = 2 a / sqrt(2)
    assert row["surface_symmetry_l2"] == pytest.approx(np.sqrt(2.0) * a, rel=1e-12)
    assert row["surface_left_max"] == pytest.approx(eta[: nx // 2].max())
    assert row["surface_right_max"] == pytest.approx(eta[nx // 2:].max())
    assert row["surface_eta_hp_rms"] == pytest.approx(a * (1 - np.cos(th)) / np.sqrt(2.0), rel=1e-10)
    # water above the bed (cells fully below 1 m are closed): sum of fills
    expected = (fill[~obs].sum()) * m.dx**3
    assert row["water_volume"] == pytest.approx(expected, rel=1e-12)
    assert row["vof_mass_total"] != row["vof_mass_total"] and row["vof_excess_count"] == -1
    # fs band: 0.5 dx above to 3 dx below the column surface on the mid plane
    assert np.isfinite(row["fs_p_rms"]) and np.isfinite(row["fs_vort_rms"])


def test_linear_wave_residual_vanishes_for_one_exact_mode():
    """A flat bed and eta = a cos(k x) with the pressure of that mode,
    rho g (base - z)^+ + rho g a cos(k x) cosh(k z') / cosh(k h): the residual
    columns are zero to rounding, and a pressure without the cosh decay is not."""
    m = fake_module(fs=True)
    nz, ny, nx = m.Nz, m.Ny, m.Nx
    a = 0.05
    kx = 2 * np.pi / m.Lx_m
    eta = a * np.cos(kx * m.x_phys)
    h = np.broadcast_to(m.free_surface_base_m + eta, (ny, nx)).copy()
    m.free_surface_h = Arr(h)
    ty = np.where(m.Z_phys < h[None], abi.FS_LIQUID, abi.FS_GAS)
    m.free_surface_type = Arr(ty)
    m.free_surface_fill = Arr((ty == abi.FS_LIQUID).astype(np.float64))
    m.ux.a[2, 1, 1] = 0.001
    depth = m.free_surface_base_m - 1.0
    zr = np.clip(m.z_phys - 1.0, 0.0, depth)
    head = np.maximum(m.free_surface_base_m - m.z_phys, 0.0)[:, None] + \
        a * np.cos(kx * m.x_phys)[None, :] * (np.cosh(kx * zr) / np.cosh(kx * depth))[:, None]
    p = m.rho_phys * m.gravity_phys * head
    m.rho.a = np.broadcast_to((1.0 + 3.0 * p / (m.rho_phys * m.vel_scale**2))[:, None, :], (nz, ny, nx)).copy()
    row = abh.frame_metrics(m, "l", 1, 0.0)
    scale = m.rho_phys * m.gravity_phys * a
    assert row["p_linear_resid_rms"] < 1e-9 * scale and row["fs_p_linear_resid_rms"] < 1e-9 * scale
    assert row["p_linear_resid_hp_rms"] < 1e-9 * scale
    assert row["p_total_rms"] > 1e-2 * scale
    undecayed = m.rho_phys * m.gravity_phys * (np.maximum(m.free_surface_base_m - m.z_phys, 0.0)[:, None]
                                               + a * np.cos(kx * m.x_phys)[None, :])
    m.rho.a = np.broadcast_to((1.0 + 3.0 * undecayed / (m.rho_phys * m.vel_scale**2))[:, None, :], (nz, ny, nx)).copy()
    assert abh.frame_metrics(m, "l", 1, 0.0)["p_linear_resid_rms"] > 1e-3 * scale


def test_vof_columns_and_counters():
    m = fake_module(fs=True)
    nz, ny, nx = m.Nz, m.Ny, m.Nx
    m.free_surface_tracking_mode = "vof"
    ty = np.full((nz, ny, nx), abi.FS_GAS, dtype=np.int32)
    fill = np.zeros((nz, ny, nx))
    ty[2:10] = abi.FS_LIQUID
    fill[2:10] = 1.0
    ty[10] = abi.FS_INTERFACE
    fill[10] = 0.3
    fill[10, 0, 0] = 0.01  # one tiny interface cell (below 0.05)
    m.free_surface_type, m.free_surface_fill = Arr(ty), Arr(fill)
    m.free_surface_h = Arr(np.full((ny, nx), 5.15))
    mass = fill * 1.01
    m.free_surface_mass = Arr(mass)
    exc = np.zeros((nz, ny, nx))
    exc[10, 1, 1] = 2e-3
    exc[0, 0, 0] = 5.0  # inside the bed: not counted
    m.free_surface_mass_excess = Arr(exc)
    counters = dict(free_surface_vof_excess_count=3, free_surface_vof_orphan_count=4,
                    free_surface_vof_collapse_cell_count=5, free_surface_vof_collapse_candidate_volume_m3=0.25,
                    free_surface_vof_collapse_applied_volume_m3=0.125, boundary_cut_link_count=11,
                    free_surface_cut_link_count=12, boundary_refill_count=13, free_surface_refill_count=14,
                    free_surface_coalescence_repair_count=15, free_surface_thin_gap_bridge_count=16,
                    free_surface_vof_detached_advect_count=17, free_surface_vof_detached_advect_mass=0.5,
                    free_surface_vof_detached_compress_count=18, free_surface_vof_detached_compress_mass=0.75)
    for k, v in counters.items():
        setattr(m, k, Cnt(v))
    m.ux.a[5, 1, 1] = 0.001
    row = abh.frame_metrics(m, "v", 1, 0.0)
    assert row["vof_mass_total"] == pytest.approx(mass[2:].sum(), rel=1e-12)
    assert (row["vof_excess_count"], row["vof_orphan_count"], row["vof_collapse_cells"]) == (3, 4, 5)
    assert (row["vof_collapse_candidate_volume"], row["vof_collapse_applied_volume"]) == (0.25, 0.125)
    assert row["vof_min_interface_fill"] == 0.01 and row["vof_tiny_interface_count"] == 1
    assert row["vof_pending_excess_mass"] == pytest.approx(2e-3) and row["vof_pending_excess_count"] == 1
    assert (row["cut_links"], row["fs_cut_links"], row["refill"], row["fs_refill"], row["fs_coalesce"],
            row["fs_bridge"]) == (11, 12, 13, 14, 15, 16)
    assert (row["vof_detached_count"], row["vof_detached_mass"], row["vof_compact_count"],
            row["vof_compact_mass"]) == (17, 0.5, 18, 0.75)
    # VOF surface band: the interface plane and the planes either side; no linear residual for VOF
    assert np.isfinite(row["fs_vort_rms"]) and np.isnan(row["fs_p_linear_resid_rms"])


def test_missing_counters_read_as_minus_one_and_nan():
    row = abh.counter_metrics(SimpleNamespace())
    assert row["cut_links"] == -1 and np.isnan(row["vof_detached_mass"])


def test_run_writes_one_row_per_frame(tmp_path, monkeypatch):
    """run() drives a module with the harness protocol: module-level steps /
    plot_freq / plot_step_start, visualize replaced, main() called."""
    src = tmp_path / "toy_solver.py"
    src.write_text(
        "import numpy as np\n"
        "import sys\n"
        f"sys.path.insert(0, {str(tmp_path)!r})\n"
        "from toy_state import build\n"
        "globals().update(build())\n"
        "steps = 0\nplot_freq = 0\nplot_step_start = 0\n"
        "def visualize(*a):\n    raise RuntimeError\n"
        "def main():\n"
        "    for s in range(plot_freq, steps + 1, plot_freq):\n"
        "        if s > plot_step_start:\n"
        "            ux.a[5, 1, 1] = 1e-3 * s\n"
        "            visualize(s, (s + 1) * 0.01, 0.0, 0.0)\n"
    )
    state = fake_module()
    (tmp_path / "toy_state.py").write_text("STATE = None\ndef build():\n    return dict(STATE.__dict__)\n")
    monkeypatch.syspath_prepend(str(tmp_path))
    import toy_state

    toy_state.STATE = state
    monkeypatch.setenv("LBM_SAVE_RENDER_SNAPSHOT", "1")
    out = abh.run(src, "toy", steps=30, plot_freq=10, plot_start=10, output_dir=tmp_path / "out")
    with out.open() as fh:
        rows = list(csv.DictReader(fh))
    assert [r["step"] for r in rows] == ["20", "30"]
    assert list(rows[0].keys()) == abh.METRIC_FIELDNAMES
    assert float(rows[1]["max_speed_phys"]) == pytest.approx(0.03 * state.vel_scale)
    snap = np.load(tmp_path / "out" / "toy_render_00030.npz")
    assert snap["obstacle"].dtype == np.uint8 and snap["water_h"].shape == (state.Ny, state.Nx)
    assert "free_surface_h" not in snap.files
