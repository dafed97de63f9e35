"""A/B harness counterpart: runs a solver module the way run_ab_case.py does
and writes the same per-frame metrics CSV.

The reference harness (run_ab_case.py) loads a solver module from a file
path (run_ab_case.py:261-279), sets steps / plot_freq / plot_step_start,
replaces visualize and calls main() (run_ab_case.py:880-912); every frame
appends one row of METRIC_FIELDNAMES (run_ab_case.py:15-36) computed from
the module's fields and counters (run_ab_case.py:286-679), and optionally
saves a render snapshot .npz (run_ab_case.py:157-201).  Its loader imports
taichi whenever a seed is given; this one passes the seed through LBM_SEED
instead, so it runs where taichi is absent.  Plots and the PyVista render
are not produced.

    python ab_harness.py --module LBM_3D_LES_air_over_waves_hip.py --label hip \\
        --steps 2000 --plot-freq 500 --output-dir out/

The column meanings follow the reference's; the functions below compute
them from the attribute surface only (fields through .to_numpy(), counters
through [None]), so any module with that surface can be measured.
"""
from __future__ import annotations

import argparse
import csv
import importlib.util
import os
import sys
from pathlib import Path

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
if _HERE not in sys.path:
    sys.path.insert(0, _HERE)

from lbm_init import linear_wave_head  # noqa: E402

# column set of the metrics file (run_ab_case.py:15-36); the order is the
# file format
METRIC_FIELDNAMES = [
    "label", "step", "time_phys", "max_speed_phys", "max_speed_x",
    "max_speed_y", "max_speed_z", "max_speed_fs_type",
    "max_speed_fs_fill", "max_speed_wet_neighbors",
    "max_speed_full_neighbors", "max_speed_gas_neighbors", "p_rms",
    "p_hp_rms", "p_total_rms", "p_total_hp_rms",
    "p_linear_resid_rms", "p_linear_resid_hp_rms", "vort_rms",
    "vort_hp_rms", "fs_p_rms", "fs_p_hp_rms", "fs_p_total_rms",
    "fs_p_total_hp_rms", "fs_p_linear_resid_rms",
    "fs_p_linear_resid_hp_rms", "fs_vort_rms", "fs_vort_hp_rms", "cut_links",
    "fs_cut_links", "refill", "fs_refill", "fs_coalesce", "fs_bridge",
    "vof_detached_count", "vof_detached_mass",
    "vof_compact_count", "vof_compact_mass",
    "surface_min", "surface_max",
    "water_volume", "vof_mass_total", "vof_excess_count",
    "vof_orphan_count", "vof_collapse_cells",
    "vof_collapse_candidate_volume", "vof_collapse_applied_volume",
    "vof_min_interface_fill", "vof_tiny_interface_count",
    "vof_pending_excess_mass", "vof_pending_excess_count",
    "surface_symmetry_l2", "surface_eta_hp_rms", "surface_left_max",
    "surface_right_max",
]

NAN = float("nan")


# ---------------------------------------------------------------------------
# small numerical helpers
def high_pass_xz(field: np.ndarray, x_periodic: bool) -> np.ndarray:
    """field minus the mean of its four (z, x) neighbours; the missing
    neighbours at the z ends (and the x ends of a walled x axis) are linear
    extrapolations, so a field linear in z or x has zero residual."""
    f = np.asarray(field, dtype=np.float64)
    zp = np.concatenate([f[1:], 2.0 * f[-1:] - f[-2:-1]], axis=0)
    zm = np.concatenate([2.0 * f[:1] - f[1:2], f[:-1]], axis=0)
    if x_periodic:
        xp, xm = np.roll(f, -1, axis=1), np.roll(f, 1, axis=1)
    else:
        xp = np.concatenate([f[:, 1:], 2.0 * f[:, -1:] - f[:, -2:-1]], axis=1)
        xm = np.concatenate([2.0 * f[:, :1] - f[:, 1:2], f[:, :-1]], axis=1)
    return f - 0.25 * (zp + zm + xp + xm)


def centred_rms(values: np.ndarray) -> float:
    """RMS about the mean, ignoring NaN."""
    return float(np.sqrt(np.nanmean((values - np.nanmean(values)) ** 2)))


def plain_rms(values: np.ndarray) -> float:
    return float(np.sqrt(np.nanmean(values**2)))


def counter(module, name: str, kind=int):
    c = getattr(module, name, None)
    if c is None:
        return -1 if kind is int else NAN
    return kind(c[None])


def linear_wave_pressure(module, eta_line, bed_line, rho_phys):
    """Pressure (Pa) of the linear finite-depth wave under the surface
    elevation eta_line on a flat bed, on the (z, x) mid plane; None for a
    non-flat bed.  The mode decay is lbm_init.linear_wave_head."""
    if not np.allclose(bed_line, bed_line[0], atol=1.0e-12, rtol=0.0):
        return None
    head = linear_wave_head(module, np.asarray(eta_line, dtype=np.float64), float(bed_line[0]))
    if getattr(module, "use_hydrostatic_balanced_pressure", 0) != 1:
        z = np.asarray(module.z_phys, dtype=np.float64)[:, None]
        head = np.maximum(float(module.free_surface_base_m) - z, 0.0) + head
    return rho_phys * float(getattr(module, "gravity_phys", 0.0)) * head


# ---------------------------------------------------------------------------
# metric groups
def surface_metrics(module, tracking, obs, fs_type, fs_fill):
    """Free-surface columns: height range, water volume, VOF mass and
    counters, interface statistics, and the shape of the y-mean surface."""
    out = {}
    h = module.free_surface_h.to_numpy()
    out["surface_min"] = float(np.nanmin(h))
    out["surface_max"] = float(np.nanmax(h))
    out["water_volume"] = float(np.nansum(np.where(obs, 0.0, fs_fill)) * module.dx**3)
    vof = tracking == "vof"
    mass = getattr(module, "free_surface_mass", None)
    out["vof_mass_total"] = float(np.nansum(np.where(obs, 0.0, mass.to_numpy()))) if (vof and mass is not None) else NAN
    out["vof_excess_count"] = counter(module, "free_surface_vof_excess_count") if vof else -1
    out["vof_collapse_cells"] = counter(module, "free_surface_vof_collapse_cell_count") if vof else -1
    out["vof_collapse_candidate_volume"] = counter(module, "free_surface_vof_collapse_candidate_volume_m3", float) if vof else NAN
    out["vof_collapse_applied_volume"] = counter(module, "free_surface_vof_collapse_applied_volume_m3", float) if vof else NAN
    if vof:
        out["vof_orphan_count"] = counter(module, "free_surface_vof_orphan_count")
        iface = (fs_type == module.FS_INTERFACE) & (~obs)
        fills = fs_fill[iface]
        out["vof_min_interface_fill"] = float(np.nanmin(fills)) if fills.size else NAN
        thr = float(getattr(module, "vof_empty_fill_threshold", 0.0))
        out["vof_tiny_interface_count"] = int(np.count_nonzero(iface & (fs_fill > 0.0) & (fs_fill < thr)))
        exc = getattr(module, "free_surface_mass_excess", None)
        if exc is not None:
            e = np.where(obs, 0.0, exc.to_numpy())
            out["vof_pending_excess_mass"] = float(np.nansum(e))
            out["vof_pending_excess_count"] = int(np.count_nonzero(np.abs(e) > 1.0e-12))
        else:
            out["vof_pending_excess_mass"], out["vof_pending_excess_count"] = NAN, -1
    else:
        out.update(vof_orphan_count=-1, vof_min_interface_fill=NAN, vof_tiny_interface_count=-1,
                   vof_pending_excess_mass=NAN, vof_pending_excess_count=-1)
    eta = h.mean(axis=0) - module.free_surface_base_m
    out["surface_symmetry_l2"] = float(np.sqrt(np.nanmean((eta - eta[::-1]) ** 2)))
    left = np.asarray(module.x_phys) < 0.5 * module.Lx_m
    out["surface_left_max"] = float(np.nanmax(eta[left]))
    out["surface_right_max"] = float(np.nanmax(eta[~left]))
    if getattr(module, "x_boundary_mode", "periodic") == "solid":
        hp = eta.copy()
        hp[1:-1] = eta[1:-1] - 0.5 * (eta[:-2] + eta[2:])
        hp[0] = eta[0] - eta[1]
        hp[-1] = eta[-1] - eta[-2]
    else:
        hp = eta - 0.5 * (np.roll(eta, 1) + np.roll(eta, -1))
    out["surface_eta_hp_rms"] = plain_rms(hp)
    return out, h


def no_surface_metrics():
    return dict(surface_min=NAN, surface_max=NAN, water_volume=NAN, vof_mass_total=NAN, vof_excess_count=-1,
                vof_orphan_count=-1, vof_collapse_cells=-1, vof_collapse_candidate_volume=NAN,
                vof_collapse_applied_volume=NAN, vof_min_interface_fill=NAN, vof_tiny_interface_count=-1,
                vof_pending_excess_mass=NAN, vof_pending_excess_count=-1, surface_symmetry_l2=NAN,
                surface_left_max=NAN, surface_right_max=NAN, surface_eta_hp_rms=NAN)


def max_speed_metrics(module, u, inactive, fs, fs_type, fs_fill, obs):
    """Location of the fastest active cell and, with a free surface, its
    type, fill and the wet / full (fill > 1/2) / gas count of its 18
    neighbours (y periodic; x periodic or walled; z walled)."""
    speed = np.sqrt(u[0] ** 2 + u[1] ** 2 + u[2] ** 2) * module.vel_scale
    speed[inactive] = np.nan
    k, j, i = np.unravel_index(np.nanargmax(speed), speed.shape)
    out = dict(max_speed_phys=float(np.nanmax(speed)), max_speed_x=float(module.x_phys[i]),
               max_speed_y=float(module.y_phys[j]), max_speed_z=float(module.z_phys[k]))
    if not fs:
        out.update(max_speed_fs_type=-1, max_speed_fs_fill=NAN, max_speed_wet_neighbors=-1,
                   max_speed_full_neighbors=-1, max_speed_gas_neighbors=-1)
        return out
    out["max_speed_fs_type"] = int(fs_type[k, j, i])
    out["max_speed_fs_fill"] = float(fs_fill[k, j, i])
    cx, cy, cz = (np.asarray(a.to_numpy()) for a in (module.cx, module.cy, module.cz))
    periodic = getattr(module, "x_boundary_mode", "periodic") == "periodic"
    wet = full = gas = 0
    for q in range(1, 19):
        ni, nj, nk = i + int(cx[q]), (j + int(cy[q])) % module.Ny, k + int(cz[q])
        if periodic:
            ni %= module.Nx
        elif not 0 <= ni < module.Nx:
            continue
        if not 0 <= nk < module.Nz or obs[nk, nj, ni]:
            continue
        if fs_type[nk, nj, ni] == module.FS_GAS:
            gas += 1
        else:
            wet += 1
            full += int(fs_fill[nk, nj, ni] > 0.5)
    out.update(max_speed_wet_neighbors=wet, max_speed_full_neighbors=full, max_speed_gas_neighbors=gas)
    return out


def band_stats(prefix, band, p, p_hp, p_tot, p_tot_hp, resid, resid_hp, vort, vort_hp):
    keys = ["p_rms", "p_hp_rms", "p_total_rms", "p_total_hp_rms", "p_linear_resid_rms", "p_linear_resid_hp_rms",
            "vort_rms", "vort_hp_rms"]
    out = {prefix + k: NAN for k in keys}
    if not np.any(band):
        return out
    out[prefix + "p_rms"] = centred_rms(p[band])
    out[prefix + "p_hp_rms"] = plain_rms(p_hp[band])
    out[prefix + "p_total_rms"] = centred_rms(p_tot[band])
    out[prefix + "p_total_hp_rms"] = plain_rms(p_tot_hp[band])
    if np.any(np.isfinite(resid[band])):
        out[prefix + "p_linear_resid_rms"] = centred_rms(resid[band])
        out[prefix + "p_linear_resid_hp_rms"] = plain_rms(resid_hp[band])
    out[prefix + "vort_rms"] = centred_rms(vort[band])
    out[prefix + "vort_hp_rms"] = plain_rms(vort_hp[band])
    return out


def midplane_metrics(module, rho, u, inactive, fs, tracking, fs_fill, fs_h, water_h):
    """Pressure and vorticity statistics on the y mid plane: within 10 cells
    of the solid boundary (0 < phi <= 10 dx), and near the free surface."""
    jm = module.Ny // 2
    rho_xz = rho[:, jm, :].astype(np.float64)
    ux_xz, uz_xz = u[0][:, jm, :].astype(np.float64), u[2][:, jm, :].astype(np.float64)
    obs_xz = inactive[:, jm, :]
    z = np.asarray(module.Z_phys)[:, jm, :]
    x_periodic = getattr(module, "x_boundary_mode", "periodic") == "periodic"
    # vorticity d(uz)/dx - d(ux)/dz: one-sided x differences at solid x
    # walls, wrapped ones otherwise (inflow/outflow included); the z
    # differences wrap (run_ab_case.py:436-444)
    if getattr(module, "x_boundary_mode", "periodic") == "solid":
        dw_dx = np.empty_like(uz_xz)
        dw_dx[:, 1:-1] = (uz_xz[:, 2:] - uz_xz[:, :-2]) / 2.0
        dw_dx[:, 0] = uz_xz[:, 1] - uz_xz[:, 0]
        dw_dx[:, -1] = uz_xz[:, -1] - uz_xz[:, -2]
    else:
        dw_dx = (np.roll(uz_xz, -1, axis=1) - np.roll(uz_xz, 1, axis=1)) / 2.0
    du_dz = (np.roll(ux_xz, -1, axis=0) - np.roll(ux_xz, 1, axis=0)) / 2.0
    vort = dw_dx - du_dz
    vort_hp = high_pass_xz(vort, x_periodic)
    # pressure: the lattice perturbation, and the total / dynamic parts the
    # pressure formulation implies
    rho_phys = getattr(module, "rho_phys", getattr(module, "rho_air_phys", 1.0))
    p_lbm = rho_phys * module.vel_scale**2 * (rho_xz - module.rho0) / 3.0
    g = float(getattr(module, "gravity_phys", 0.0))
    surface_z = None
    if fs:
        column = tracking != "vof"
        if getattr(module, "use_component_balanced_pressure", 0) == 1:
            ref = getattr(module, "pressure_reference_head", None)
            ref_xz = ref.to_numpy()[:, jm, :] if ref is not None else np.zeros_like(p_lbm)
            p_tot = p_lbm + rho_phys * g * ref_xz
            p, p_cmp = p_tot, p_tot
        elif getattr(module, "use_hydrostatic_balanced_pressure", 0) == 1:
            p_tot = p_lbm + rho_phys * g * (float(module.free_surface_base_m) - z)
            if column:
                surface_z = fs_h[jm, :][None, :]
                p, p_cmp = p_lbm, p_lbm
            else:
                p, p_cmp = p_tot, p_tot
        else:
            p_tot = p_lbm
            if column:
                surface_z = fs_h[jm, :][None, :]
                p = p_tot - rho_phys * g * np.maximum(surface_z - z, 0.0)
            else:
                p = p_tot
            p_cmp = p_tot
    else:
        p_tot, p, p_cmp = p_lbm, p_lbm, p_lbm
    p_hp, p_tot_hp = high_pass_xz(p, x_periodic), high_pass_xz(p_tot, x_periodic)
    resid = np.full_like(p, np.nan)
    resid_hp = np.full_like(p, np.nan)
    if fs and tracking != "vof":
        expected = linear_wave_pressure(module, fs_h[jm, :] - module.free_surface_base_m, water_h[jm, :], rho_phys)
        if expected is not None:
            resid = p_cmp - expected
            resid_hp = high_pass_xz(resid, x_periodic)
    phi_xz = module.phi.to_numpy()[:, jm, :]
    band = (phi_xz > 0.0) & (phi_xz <= 10.0 * module.dx) & (~obs_xz)
    out = band_stats("", band, p, p_hp, p_tot, p_tot_hp, resid, resid_hp, vort, vort_hp)
    fs_band = np.zeros_like(band)
    if fs:
        if tracking == "vof":
            fill_xz = fs_fill[:, jm, :]
            seed = (fill_xz > 0.0) & (fill_xz < 1.0) & (~obs_xz)
            fs_band = seed.copy()
            fs_band[1:, :] |= seed[:-1, :]
            fs_band[:-1, :] |= seed[1:, :]
            fs_band[:, 1:] |= seed[:, :-1]
            fs_band[:, :-1] |= seed[:, 1:]
            fs_band &= ~obs_xz
        else:
            surf = fs_h[jm, :][None, :]
            fs_band = (z <= surf + 0.5 * module.dx) & (z >= surf - 3.0 * module.dx) & (~obs_xz)
    out.update(band_stats("fs_", fs_band, p, p_hp, p_tot, p_tot_hp, resid, resid_hp, vort, vort_hp))
    return out


def counter_metrics(module):
    return {
        "cut_links": counter(module, "boundary_cut_link_count"),
        "fs_cut_links": counter(module, "free_surface_cut_link_count"),
        "refill": counter(module, "boundary_refill_count"),
        "fs_refill": counter(module, "free_surface_refill_count"),
        "fs_coalesce": counter(module, "free_surface_coalescence_repair_count"),
        "fs_bridge": counter(module, "free_surface_thin_gap_bridge_count"),
        "vof_detached_count": counter(module, "free_surface_vof_detached_advect_count"),
        "vof_detached_mass": counter(module, "free_surface_vof_detached_advect_mass", float),
        "vof_compact_count": counter(module, "free_surface_vof_detached_compress_count"),
        "vof_compact_mass": counter(module, "free_surface_vof_detached_compress_mass", float),
    }


def frame_metrics(module, label: str, step: int, time_phys: float) -> dict:
    """One row of the metrics file (run_ab_case.py:286-679)."""
    rho = module.rho.to_numpy()
    u = [module.ux.to_numpy(), module.uy.to_numpy(), module.uz.to_numpy()]
    obs = module.obstacle.to_numpy().astype(bool)
    water_h = module.water_h.to_numpy()
    fs = getattr(module, "water_free_surface_enabled", 0) == 1
    tracking = getattr(module, "free_surface_tracking_mode", "height")
    inactive = obs.copy()
    fs_type = fs_fill = None
    fs_h = water_h
    row = {"label": label, "step": step, "time_phys": time_phys}
    if fs:
        fs_type = module.free_surface_type.to_numpy()
        fs_fill = module.free_surface_fill.to_numpy()
        inactive |= (fs_type == module.FS_GAS) & (~obs)
        surf, fs_h = surface_metrics(module, tracking, obs, fs_type, fs_fill)
        row.update(surf)
    else:
        row.update(no_surface_metrics())
    row.update(max_speed_metrics(module, u, inactive, fs, fs_type, fs_fill, obs))
    row.update(midplane_metrics(module, rho, u, inactive, fs, tracking, fs_fill, fs_h, water_h))
    row.update(counter_metrics(module))
    return row


def save_render_snapshot(module, label, step, time_phys, output_dir: Path):
    """The render snapshot (run_ab_case.py:157-201) when
    LBM_SAVE_RENDER_SNAPSHOT is set: grid, bed, obstacle, surface, fill."""
    if os.environ.get("LBM_SAVE_RENDER_SNAPSHOT", "").strip().lower() not in ("1", "true", "yes", "on"):
        return None
    fs = getattr(module, "water_free_surface_enabled", 0) == 1
    m = module
    data = {
        "time_phys": np.array(time_phys, dtype=np.float64),
        "dx": np.array(m.dx, dtype=np.float64),
        "lx_m": np.array(m.Lx_m, dtype=np.float64),
        "ly_m": np.array(m.Ly_m, dtype=np.float64),
        "lz_m": np.array(m.Lz_m, dtype=np.float64),
        "x_phys": np.asarray(m.x_phys, dtype=np.float32),
        "y_phys": np.asarray(m.y_phys, dtype=np.float32),
        "z_phys": np.asarray(m.z_phys, dtype=np.float32),
        "water_h": m.water_h.to_numpy().astype(np.float32),
        "obstacle": m.obstacle.to_numpy().astype(np.uint8),
        "free_surface_enabled": np.array(int(fs), dtype=np.int8),
        "tracking_mode": np.array(str(getattr(m, "free_surface_tracking_mode", "height"))),
        "include_cube_geometry": np.array(int(getattr(m, "include_cube_geometry", 0)), dtype=np.int8),
        "cube_size_m": np.array([float(getattr(m, f"cube_size_{a}_m", getattr(m, "cube_side_m", 0.0))) for a in "xyz"],
                                dtype=np.float32),
        "cube_center_m": np.array([float(getattr(m, f"cube_center_{a}_m", 0.0)) for a in "xyz"], dtype=np.float32),
        "cube_yaw_pitch_deg": np.array([float(getattr(m, "cube_yaw_deg", 0.0)), float(getattr(m, "cube_pitch_deg", 0.0))],
                                       dtype=np.float32),
    }
    if fs:
        data["free_surface_h"] = m.free_surface_h.to_numpy().astype(np.float32)
        data["free_surface_fill"] = m.free_surface_fill.to_numpy().astype(np.float32)
    path = output_dir / f"{label}_render_{step:05d}.npz"
    np.savez_compressed(path, **data)
    return path


def load_module(path: Path, label: str, seed=None):
    """Import a solver module from a file path (run_ab_case.py:261-279); the
    seed reaches the module through LBM_SEED unless that is already set."""
    if seed is not None:
        os.environ.setdefault("LBM_SEED", str(int(seed)))
    spec = importlib.util.spec_from_file_location(f"lbm_ab_{label}", path)
    if spec is None or spec.loader is None:
        raise RuntimeError(f"Could not load solver from {path}")
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


def run(module_path: Path, label: str, steps: int, plot_freq: int, plot_start: int, output_dir: Path, seed=None) -> Path:
    output_dir.mkdir(parents=True, exist_ok=True)
    metrics_path = output_dir / f"{label}_metrics.csv"
    module = load_module(module_path, label, seed)
    module.steps = steps
    module.plot_freq = plot_freq
    module.plot_step_start = plot_start
    with metrics_path.open("w", newline="") as fh:
        writer = csv.DictWriter(fh, fieldnames=METRIC_FIELDNAMES)
        writer.writeheader()

        def visualize(step, time_phys, re_phys, re_lb):
            row = frame_metrics(module, label, step, time_phys)
            writer.writerow(row)
            fh.flush()
            snap = save_render_snapshot(module, label, step, time_phys, output_dir)
            note = f", render_snapshot={snap}" if snap is not None else ""
            print(f"[{label}] step {step}: max_speed={row['max_speed_phys']:.3e} m/s, p_hp_rms={row['p_hp_rms']:.3e} Pa, "
                  f"vort_hp_rms={row['vort_hp_rms']:.3e}, cut_links={row['cut_links']}, "
                  f"surface=[{row['surface_min']:.3f},{row['surface_max']:.3f}] m, volume={row['water_volume']:.3f} m^3"
                  f"{note}", flush=True)

        module.visualize = visualize
        print(f"[{label}] Starting {steps} steps from {module_path.name}; metrics -> {metrics_path}", flush=True)
        module.main()
    if hasattr(module, "close"):
        module.close()
    return metrics_path


def main(argv=None):
    parser = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    parser.add_argument("--module", required=True)
    parser.add_argument("--label", required=True)
    parser.add_argument("--steps", type=int, default=10000)
    parser.add_argument("--plot-freq", type=int, default=500)
    parser.add_argument("--plot-start", type=int, default=0)
    parser.add_argument("--output-dir", required=True)
    parser.add_argument("--random-seed", type=int, default=12345)
    args = parser.parse_args(argv)
    run(Path(args.module).resolve(), args.label, args.steps, args.plot_freq, args.plot_start,
        Path(args.output_dir).resolve(), args.random_seed)


if __name__ == "__main__":
    main()
