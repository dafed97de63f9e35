"""LBM_* environment options -> run configuration.

Restates the configuration block of the reference solver
(LBM_3D_LES_air_over_waves_CODEX.py:35-529, "CODEX" below) name for name:
same variables, same defaults, same validation errors, same derived
constants (vel_scale, dt_phys, g_lb, tau0, ...).  The result is a
``SolverConfig`` whose attributes carry the reference's module-level names
(Nx, dx, vel_scale, ...) and whose ``to_struct()`` gives the ``lbm_config_t``
the C ABI takes.

New options (absent from the reference; their defaults keep reference
behaviour):

  LBM_TAU                 relaxation time set directly (else CODEX:463-466)
  LBM_LES_CS              Smagorinsky constant (reference: fixed 0.2, CODEX:494)
  LBM_G_LB                lattice gravity set directly (else CODEX:451)
  LBM_TOP_BOUNDARY        keep (reference) | free_slip
  LBM_X_BOUNDARY          adds inflow_outflow to periodic | solid
  LBM_Y_BOUNDARY          periodic (reference) | solid
  LBM_BED_PROFILE         adds moving_sine1: eta = A sin(k(x - c t))
  LBM_BED_PHASE_SPEED_LU  phase speed c of moving_sine1, lattice units
  LBM_STORAGE             fp32 | fp16 | bf16 | fp64 population storage
  LBM_STORAGE_SHIFT       k of the (f_i - s_i) * 2^k scaling (fp16/bf16);
                          default chosen from a bound of max |f_i - s_i|
  LBM_STORAGE_U_REF       auto | 0 | "ux,uy,uz": velocity of the shift
                          s_i = feq_i(1, u_ref) (0 gives s_i = w_i)
  LBM_STORAGE_ROUNDING    stochastic (default) | nearest
  LBM_KERNEL_PATH         auto | staged | fused
  LBM_INITIAL_STATE       reference | uniform_noise | loglaw_noise
  LBM_INIT_U_LU, LBM_INIT_NOISE_LU     uniform_noise mean and half-width
  LBM_INFLOW_U_LU, LBM_INFLOW_Z0_CELLS, LBM_INFLOW_NOISE  log-law inflow
  LBM_SEED                seed of all host-side draws and of backscatter
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Mapping, Optional

import numpy as np

import lbm_abi as abi


def env_choice(env: Mapping[str, str], name: str, default: str, choices: tuple) -> str:
    """lbm_env.py:9-14: lower-cased value, ValueError when not allowed."""
    value = str(env.get(name, default)).lower()
    if value not in choices:
        allowed = ", ".join(repr(choice) for choice in choices)
        raise ValueError(f"{name} must be one of: {allowed}")
    return value


def env_int(env: Mapping[str, str], name: str, default: int) -> int:
    """lbm_env.py:17-18."""
    return int(env.get(name, default))


def _flag(env, name, default):
    return str(env.get(name, default)).lower() in ("1", "true", "yes")


@dataclass
class SolverConfig:
    """Every derived constant of CODEX:35-529 the step needs, under the
    reference's names, plus the new options."""

    values: dict = field(default_factory=dict)

    def __getattr__(self, name):
        try:
            return self.__dict__["values"][name]
        except KeyError as exc:
            raise AttributeError(name) from exc

    def init_params(self) -> abi.LbmInitParams:
        """lbm_init_params_t of this configuration (the host initial fields,
        CODEX:2714-2800, 1197-1287, 743-818)."""
        v = self.values
        p = abi.LbmInitParams()
        p.version = abi.LBM_INIT_PARAMS_VERSION
        p.initial_state = {"reference": abi.INIT_REFERENCE, "uniform_noise": abi.INIT_UNIFORM_NOISE,
                           "loglaw_noise": abi.INIT_LOGLAW_NOISE}[v["initial_state"]]
        p.fs_initial = {"gaussian": abi.FSI_GAUSSIAN, "solitary": abi.FSI_SOLITARY}.get(
            v["free_surface_initial_condition"], abi.FSI_FLAT)
        p.vof_block = int(v["vof_initial_shape"] == "block")
        p.linear_wave_pressure = int(v["initial_pressure_mode"] == "linear_wave")
        p.solitary_velocity = {"none": abi.SOLVEL_NONE, "shallow_water": abi.SOLVEL_SHALLOW,
                               "shallow_water_divfree": abi.SOLVEL_SHALLOW_DIVFREE}[v["solitary_initial_velocity_mode"]]
        p.hydrostatic_balanced = int(v["use_hydrostatic_balanced_pressure"])
        p.component_balanced = int(v["use_component_balanced_pressure"])
        seed = int(v["seed"])
        if not 0 <= seed < 1 << 64:
            raise ValueError(f"LBM_SEED={seed}: the host draws take a seed in [0, 2**64)")
        p.seed = seed
        for name in ("dx", "offset_m", "free_surface_base_m", "free_surface_depth_m", "rho0", "g_lb", "u_current_lb",
                     "u_top_lb", "tau0", "kappa", "vel_scale", "init_u_lu", "init_noise_lu", "inflow_u_lu",
                     "inflow_z0_cells", "inflow_noise", "gaussian_amp_m", "gaussian_sigma_m", "gaussian_center_x_m",
                     "solitary_amp_m", "solitary_x0_m", "solitary_direction", "solitary_celerity_phys",
                     "solitary_kappa_m_inv"):
            setattr(p, name, float(v[name]))
        p.turbulence_intensity = float(v["TI"])
        for a, ax in enumerate("xyz"):
            p.vof_block_size_m[a] = float(v[f"vof_block_size_{ax}_m"])
            p.vof_block_center_m[a] = float(v[f"vof_block_center_{ax}_m"])
        return p

    def to_struct(self) -> abi.LbmConfig:
        v = self.values
        s = abi.LbmConfig()
        s.version = abi.LBM_CONFIG_VERSION
        s.nx, s.ny, s.nz = v["Nx"], v["Ny"], v["Nz"]
        s.physics_mode = abi.PHYS_WATER if v["physics_mode"] == "water" else abi.PHYS_AIR
        s.collision = abi.COLL_REGULARIZED if v["collision_model"] == "regularized" else abi.COLL_KBC
        s.use_les = 1 if v["use_LES"] else 0
        s.x_boundary = {"periodic": abi.XB_PERIODIC, "solid": abi.XB_SOLID, "inflow_outflow": abi.XB_INFLOW_OUTFLOW}[
            v["x_boundary_mode"]
        ]
        s.y_boundary = abi.YB_SOLID if v["y_boundary_mode"] == "solid" else abi.YB_PERIODIC
        s.top_boundary = abi.TOP_FREE_SLIP if v["top_boundary_mode"] == "free_slip" else abi.TOP_KEEP
        s.bed_profile = {
            "moving_sine": abi.BED_MOVING_SINE,
            "flat": abi.BED_FLAT,
            "piecewise_linear": abi.BED_PIECEWISE,
            "moving_sine1": abi.BED_MOVING_SINE1,
        }[v["bed_profile"]]
        s.include_bed = v["include_bed_geometry"]
        s.include_cube = v["include_cube_geometry"]
        s.phi_fraction = 1 if v["boundary_geometry_mode"] in ("phi_fraction", "phi") else 0
        s.phi_flags = 1 if v["boundary_geometry_mode"] == "phi" else 0
        s.init_fs_populations = 1 if v["initialize_free_surface_populations"] else 0
        s.solitary[:] = [v["solitary_amp_m"], v["free_surface_depth_m"], v["solitary_x0_m"], v["solitary_direction"],
                         v["solitary_celerity_phys"]]
        s.cube_half[:] = v["cube_half"]
        s.cube_center[:] = v["cube_center"]
        s.cube_rot[:] = v["cube_rot"]
        s.wall_velocity_link = v["use_link_wall_velocity"]
        s.enable_top_drive = v["enable_top_drive"]
        s.enable_x_fringe = v["enable_x_fringe"]
        s.enable_top_sponge = v["enable_top_sponge"]
        s.storage = {"fp32": abi.STORE_FP32, "fp16": abi.STORE_FP16, "bf16": abi.STORE_BF16, "fp64": abi.STORE_FP64}[
            v["storage"]
        ]
        s.storage_shift = v["storage_shift"]
        s.storage_rounding = abi.ROUND_STOCHASTIC if v["storage_rounding"] == "stochastic" else abi.ROUND_NEAREST
        for a in range(3):
            s.storage_u_ref[a] = v["storage_u_ref"][a]
        s.kernel_path = {"auto": abi.PATH_AUTO, "staged": abi.PATH_STAGED, "fused": abi.PATH_FUSED}[v["kernel_path"]]
        s.fs_enabled = v["water_free_surface_enabled"]
        s.fs_mode = {"none": abi.FSM_NONE, "height_static": abi.FSM_HEIGHT_STATIC,
                     "height_kinematic": abi.FSM_HEIGHT_KINEMATIC,
                     "prescribed_solitary": abi.FSM_PRESCRIBED_SOLITARY}[v["free_surface_mode"]]
        s.fs_tracking = abi.FST_VOF if v["free_surface_tracking_mode"] == "vof" else abi.FST_HEIGHT
        s.pressure_formulation = {
            "total_pressure": abi.PF_TOTAL,
            "hydrostatic_balanced": abi.PF_HYDROSTATIC_BALANCED,
            "vof_component_balanced": abi.PF_VOF_COMPONENT_BALANCED,
        }[v["pressure_formulation"]]
        s.fs_boundary_mode = abi.FSB_HYDROSTATIC if v["free_surface_boundary_mode"] == "hydrostatic" else abi.FSB_CELL
        s.refill_mode = {
            "equilibrium": abi.REFILL_EQUILIBRIUM,
            "noneq": abi.REFILL_NONEQ,
            "directional": abi.REFILL_DIRECTIONAL,
        }[v["free_surface_refill_mode"]]
        s.vof_link_aperture = v["vof_link_aperture_enabled"]
        s.vof_thin_gap_bridge = 1 if v["vof_thin_gap_bridge_enabled"] else 0
        s.vof_detached_advection = 1 if v["vof_detached_advection_enabled"] else 0
        s.vof_collapse_trapped_gas = 1 if v["vof_collapse_trapped_gas_enabled"] else 0
        s.vof_orphan_search_radius = v["vof_orphan_search_radius"]
        s.vof_orphan_max_weak_neighbors = v["vof_orphan_max_weak_neighbors"]
        s.vof_detached_max_wet_neighbors = v["vof_detached_max_wet_neighbors"]
        s.vof_detached_max_resolved_neighbors = v["vof_detached_max_resolved_neighbors"]
        s.vof_collapse_interval = v["vof_collapse_interval"]
        s.vof_collapse_flood_sweeps = v["vof_collapse_flood_sweeps"]
        s.seed = v["seed"] & 0xFFFFFFFF
        s.dx = v["dx"]
        s.rho0 = v["rho0"]
        s.tau0 = v["tau0"]
        s.nu0 = v["nu0"]
        s.cs_les = v["Cs"]
        s.delta_les = v["Delta"]
        s.nu_t_max = v["nu_t_max"]
        s.kappa = v["kappa"]
        s.z0_phys = v["z0_phys"]
        s.u_top_lb = v["u_top_lb"]
        s.u_current_lb = v["u_current_lb"]
        s.g_lb = v["g_lb"]
        s.g_body_lb = v["g_body_lb"]
        s.c_backscatter = v["C_backscatter"]
        s.offset_m = v["offset_m"]
        s.bed_amp_m = v["bed_amp_m"]
        s.bed_wavelength_m = v["bed_wavelength_m"]
        s.wave_speed = v["wave_speed"]
        s.wave_period = v["wave_period"]
        s.depth_swl_m = v["depth_swl_m"]
        s.vel_scale = v["vel_scale"]
        s.bed_phase_speed_mps = v["bed_phase_speed_mps"]
        s.free_surface_base_m = v["free_surface_base_m"]
        s.vof_empty_fill_threshold = v["vof_empty_fill_threshold"]
        s.vof_orphan_empty_threshold = v["vof_orphan_empty_threshold"]
        s.vof_detached_residual_fill_threshold = v["vof_detached_residual_fill_threshold"]
        s.vof_thin_gap_bridge_strength = v["vof_thin_gap_bridge_strength"]
        s.vof_collapse_max_volume_m3 = v["vof_collapse_max_volume_m3"]
        s.dt_phys = v["dt_phys"]
        return s


def load_piecewise_bottom_profile(path: str, Nx: int, dx: float, Lx_m: float):
    """CODEX:223-276: `x z` pairs -> bed height at the x cell centres."""
    resolved = os.path.abspath(path)
    if not os.path.exists(resolved):
        raise FileNotFoundError(f"LBM_BOTTOM_FILE not found: {resolved}")
    pairs = []
    with open(resolved, "r", encoding="utf-8") as fh:
        for line_number, raw in enumerate(fh, start=1):
            line = raw.split("#", 1)[0].replace(",", " ").strip()
            if not line:
                continue
            parts = line.split()
            if len(parts) != 2:
                raise ValueError(
                    f"Invalid bottom profile line {line_number} in {resolved}: expected exactly two values: x_m z_m"
                )
            try:
                xv, zv = float(parts[0]), float(parts[1])
            except ValueError as exc:
                raise ValueError(f"Invalid numeric value on bottom profile line {line_number} in {resolved}") from exc
            if not (np.isfinite(xv) and np.isfinite(zv)):
                raise ValueError(f"Non-finite bottom profile value on line {line_number} in {resolved}")
            pairs.append((xv, zv))
    if len(pairs) < 2:
        raise ValueError(f"LBM_BOTTOM_FILE must contain at least two x z pairs: {resolved}")
    pairs.sort(key=lambda item: item[0])
    xs = np.array([p[0] for p in pairs], dtype=np.float64)
    zs = np.array([p[1] for p in pairs], dtype=np.float64)
    if np.any(np.diff(xs) <= 0.0):
        raise ValueError(f"LBM_BOTTOM_FILE contains duplicate x coordinates: {resolved}")
    x_centers = (np.arange(Nx, dtype=np.float64) + 0.5) * dx
    z_centers = np.interp(x_centers, xs, zs, left=zs[0], right=zs[-1]).astype(np.float32)
    meta = {
        "path": resolved,
        "point_count": len(pairs),
        "first_x": float(xs[0]),
        "last_x": float(xs[-1]),
        "min_z": float(np.min(z_centers)),
        "max_z": float(np.max(z_centers)),
        "z_at_lx": float(np.interp(Lx_m, xs, zs, left=zs[0], right=zs[-1])),
    }
    return z_centers, meta


def parse_env(env: Optional[Mapping[str, str]] = None) -> SolverConfig:
    """CODEX:35-529 (same order, defaults and errors) + the new options."""
    env = os.environ if env is None else env
    v: dict = {}
    physics_mode = env_choice(env, "LBM_PHYSICS_MODE", "air", ("air", "water"))
    free_surface_mode = str(env.get("LBM_FREE_SURFACE_MODE", "none" if physics_mode == "air" else "height_kinematic")).lower()
    if free_surface_mode not in ("none", "height_static", "height_kinematic", "prescribed_solitary"):
        raise ValueError(
            "LBM_FREE_SURFACE_MODE must be 'none', 'height_static', 'height_kinematic', or 'prescribed_solitary'"
        )
    water_free_surface_enabled = 1 if free_surface_mode != "none" else 0
    fs_initial = str(
        env.get(
            "LBM_FREE_SURFACE_INITIAL",
            "solitary"
            if (physics_mode == "water" and free_surface_mode in ("height_kinematic", "prescribed_solitary"))
            else "flat",
        )
    ).lower()
    if fs_initial not in ("flat", "solitary", "gaussian"):
        raise ValueError("LBM_FREE_SURFACE_INITIAL must be 'flat', 'solitary', or 'gaussian'")
    vof_initial_shape = str(env.get("LBM_VOF_INITIAL_SHAPE", "none")).lower()
    if vof_initial_shape not in ("none", "block"):
        raise ValueError("LBM_VOF_INITIAL_SHAPE must be 'none' or 'block'")
    initial_pressure_mode = env_choice(env, "LBM_INITIAL_PRESSURE_MODE", "hydrostatic", ("hydrostatic", "linear_wave"))
    default_pf = "hydrostatic_balanced" if physics_mode == "water" else "total_pressure"
    if physics_mode == "water" and vof_initial_shape != "none":
        default_pf = "vof_component_balanced"
    pressure_formulation = env_choice(
        env,
        "LBM_PRESSURE_FORMULATION",
        default_pf,
        ("total_pressure", "hydrostatic_balanced", "vof_component_balanced"),
    )
    pf_auto = False
    tracking = str(env.get("LBM_FREE_SURFACE_TRACKING", "height")).lower()
    if tracking not in ("height", "vof"):
        raise ValueError("LBM_FREE_SURFACE_TRACKING must be 'height' or 'vof'")
    fs_boundary = str(env.get("LBM_FREE_SURFACE_BOUNDARY", "cell")).lower()
    if fs_boundary not in ("cell", "hydrostatic"):
        raise ValueError("LBM_FREE_SURFACE_BOUNDARY must be 'cell' or 'hydrostatic'")
    refill = str(env.get("LBM_FREE_SURFACE_REFILL", "equilibrium")).lower()
    if refill not in ("equilibrium", "noneq", "directional"):
        raise ValueError("LBM_FREE_SURFACE_REFILL must be 'equilibrium', 'noneq', or 'directional'")
    init_fs_pops = _flag(env, "LBM_INIT_FREE_SURFACE_POPULATIONS", "0")
    vof_link_aperture_enabled = 1 if _flag(env, "LBM_VOF_LINK_APERTURE", "1") else 0
    vof_empty = float(env.get("LBM_VOF_EMPTY_FILL_THRESHOLD", "0.0"))
    if not (0.0 <= vof_empty < 0.5):
        raise ValueError("LBM_VOF_EMPTY_FILL_THRESHOLD must be in [0, 0.5)")
    vof_orphan_thr = float(env.get("LBM_VOF_ORPHAN_EMPTY_THRESHOLD", "0.49"))
    if not (0.0 <= vof_orphan_thr < 0.5):
        raise ValueError("LBM_VOF_ORPHAN_EMPTY_THRESHOLD must be in [0, 0.5)")
    vof_orphan_r = int(env.get("LBM_VOF_ORPHAN_SEARCH_RADIUS", "4"))
    if vof_orphan_r < 1 or vof_orphan_r > 4:
        raise ValueError("LBM_VOF_ORPHAN_SEARCH_RADIUS must be in [1, 4]")
    vof_orphan_weak = int(env.get("LBM_VOF_ORPHAN_MAX_WEAK_NEIGHBORS", "1"))
    if vof_orphan_weak < 0 or vof_orphan_weak > 18:
        raise ValueError("LBM_VOF_ORPHAN_MAX_WEAK_NEIGHBORS must be in [0, 18]")
    vof_det_adv = _flag(env, "LBM_VOF_DETACHED_ADVECTION", "1")
    vof_det_wet = int(env.get("LBM_VOF_DETACHED_MAX_WET_NEIGHBORS", "18"))
    if vof_det_wet < 0 or vof_det_wet > 18:
        raise ValueError("LBM_VOF_DETACHED_MAX_WET_NEIGHBORS must be in [0, 18]")
    vof_det_res = int(env.get("LBM_VOF_DETACHED_MAX_RESOLVED_NEIGHBORS", "0"))
    if vof_det_res < 0 or vof_det_res > 18:
        raise ValueError("LBM_VOF_DETACHED_MAX_RESOLVED_NEIGHBORS must be in [0, 18]")
    vof_det_resid = float(env.get("LBM_VOF_DETACHED_RESIDUAL_FILL_THRESHOLD", "0.01"))
    if not (0.0 <= vof_det_resid < 0.5):
        raise ValueError("LBM_VOF_DETACHED_RESIDUAL_FILL_THRESHOLD must be in [0, 0.5)")
    vof_bridge = _flag(env, "LBM_VOF_THIN_GAP_BRIDGE", "0")
    vof_bridge_strength = float(env.get("LBM_VOF_THIN_GAP_BRIDGE_STRENGTH", "1.0"))
    if vof_bridge_strength < 0.0:
        raise ValueError("LBM_VOF_THIN_GAP_BRIDGE_STRENGTH must be nonnegative")
    vof_collapse = _flag(env, "LBM_VOF_COLLAPSE_TRAPPED_GAS", "0")

    U_top_phys = 30.0 if physics_mode == "air" else 0.0
    U_top_phys = float(env.get("LBM_U_TOP_MPS", U_top_phys))
    water_current_phys = float(
        env.get("LBM_WATER_CURRENT_MPS", 0.0 if physics_mode == "water" else U_top_phys)
    )
    Lx_m = float(env.get("LBM_LX_M", 300.0))
    Ly_m = float(env.get("LBM_LY_M", 10.0))
    Lz_m = float(env.get("LBM_LZ_M", 20.0))
    dx = float(env.get("LBM_DX_M", 0.5))
    nu_air = 1.516e-5
    rho_air_standard_phys = 1.2
    nu_water = float(env.get("LBM_WATER_NU_M2S", 1.004e-6))
    rho_water_phys = float(env.get("LBM_WATER_RHO_KGM3", 998.2))
    fluid_name = "water" if physics_mode == "water" else "air"
    nu_phys = nu_water if physics_mode == "water" else nu_air
    rho_phys = rho_water_phys if physics_mode == "water" else rho_air_standard_phys
    gravity_phys = float(env.get("LBM_GRAVITY_MPS2", 9.81 if physics_mode == "water" else 0.0))
    Nx = int(Lx_m / dx)
    Ny = int(Ly_m / dx)
    Nz = int(Lz_m / dx)
    vof_collapse_interval = env_int(env, "LBM_VOF_COLLAPSE_INTERVAL", 200)
    if vof_collapse_interval < 1:
        raise ValueError("LBM_VOF_COLLAPSE_INTERVAL must be positive")
    vof_collapse_max_volume_m3 = float(env.get("LBM_VOF_COLLAPSE_MAX_VOLUME_M3", "8.0"))
    if vof_collapse_max_volume_m3 < 0.0:
        raise ValueError("LBM_VOF_COLLAPSE_MAX_VOLUME_M3 must be nonnegative")
    vof_collapse_flood_sweeps = env_int(env, "LBM_VOF_COLLAPSE_FLOOD_SWEEPS", max(1, Nx + Ny + Nz))
    if vof_collapse_flood_sweeps < 1:
        raise ValueError("LBM_VOF_COLLAPSE_FLOOD_SWEEPS must be positive")
    if vof_collapse_flood_sweeps < Nx + Ny + Nz:
        # the reference's sweeps update the marks in place in parallel, so a
        # fill cut short depends on the sweep schedule; this build runs the
        # flood fill to its fixed point (exact 6-connected reachability)
        raise ValueError(
            "LBM_VOF_COLLAPSE_FLOOD_SWEEPS below Nx + Ny + Nz is not supported: the trapped-gas flood fill runs to "
            "its fixed point here (a partial fill depends on the sweep schedule); leave it unset or use a value "
            ">= Nx + Ny + Nz, which bounds the sweeps"
        )
    steps = env_int(env, "LBM_STEPS", 400000)
    plot_freq = env_int(env, "LBM_PLOT_FREQ", 100)
    plot_step_start = env_int(env, "LBM_PLOT_START", 0)
    boundary_geometry_mode = str(env.get("LBM_BOUNDARY_GEOMETRY", "vertical")).lower()
    if boundary_geometry_mode not in ("vertical", "phi_fraction", "phi"):
        raise ValueError("LBM_BOUNDARY_GEOMETRY must be 'vertical', 'phi_fraction', or 'phi'")
    default_xb = "solid" if (physics_mode == "water" and fs_initial == "gaussian") else "periodic"
    x_boundary_mode = env_choice(env, "LBM_X_BOUNDARY", default_xb, ("periodic", "solid", "inflow_outflow"))
    y_boundary_mode = env_choice(env, "LBM_Y_BOUNDARY", "periodic", ("periodic", "solid"))
    wall_velocity_sampling = env_choice(env, "LBM_WALL_VELOCITY_SAMPLING", "link", ("cell", "link"))
    use_link_wall_velocity = 1 if wall_velocity_sampling == "link" else 0

    bed_profile = str(env.get("LBM_BED_PROFILE", "moving_sine" if physics_mode == "air" else "flat")).lower()
    if bed_profile not in ("moving_sine", "flat", "piecewise_linear", "moving_sine1"):
        raise ValueError("LBM_BED_PROFILE must be 'moving_sine', 'flat', 'piecewise_linear', or 'moving_sine1'")
    bed_is_moving = 1 if bed_profile in ("moving_sine", "moving_sine1") else 0
    bottom_file_path = env.get("LBM_BOTTOM_FILE", "bottom.txt")
    bed_amp_m = 1.5 if bed_profile in ("moving_sine", "moving_sine1") else 0.0
    bed_amp_m = float(env.get("LBM_BED_AMP_M", bed_amp_m))
    depth_swl_m = float(env.get("LBM_WATER_DEPTH_M", 10.0))
    wave_speed = float(np.sqrt(gravity_phys * depth_swl_m) if gravity_phys > 0.0 else np.sqrt(9.81 * depth_swl_m))
    bed_wavelength_m = float(env.get("LBM_BED_WAVELENGTH_M", 30.0))
    wave_period = bed_wavelength_m / max(wave_speed, 1e-12)
    offset_m = 2.0 * bed_amp_m + 3.0 * dx if bed_profile in ("moving_sine", "moving_sine1") else 0.0
    offset_m = float(env.get("LBM_BED_LEVEL_M", offset_m))
    # moving_sine1 prescribes its phase speed in lattice units, so it does not
    # enter the velocity scale
    U_wave_phys = wave_speed if bed_profile == "moving_sine" else 0.0

    if bed_profile == "piecewise_linear":
        piecewise_bed_z_np, piecewise_meta = load_piecewise_bottom_profile(bottom_file_path, Nx, dx, Lx_m)
    else:
        piecewise_bed_z_np = np.full(Nx, offset_m, dtype=np.float32)
        piecewise_meta = {
            "path": "",
            "point_count": 0,
            "first_x": 0.0,
            "last_x": Lx_m,
            "min_z": float(offset_m),
            "max_z": float(offset_m),
            "z_at_lx": float(offset_m),
        }

    free_surface_depth_m = float(env.get("LBM_FREE_SURFACE_DEPTH_M", depth_swl_m))
    free_surface_base_m = float(env.get("LBM_FREE_SURFACE_LEVEL_M", offset_m + free_surface_depth_m))
    if free_surface_depth_m <= 0.0:
        raise ValueError("LBM_FREE_SURFACE_DEPTH_M must be positive")
    sol_ratio_env = env.get("LBM_SOLITARY_HEIGHT_DEPTH_RATIO")
    if sol_ratio_env is not None:
        solitary_height_depth_ratio = float(sol_ratio_env)
        if solitary_height_depth_ratio <= 0.0:
            raise ValueError("LBM_SOLITARY_HEIGHT_DEPTH_RATIO must be positive")
        solitary_amp_m = solitary_height_depth_ratio * free_surface_depth_m
    else:
        solitary_amp_m = float(env.get("LBM_SOLITARY_AMPLITUDE_M", 0.5))
        if solitary_amp_m <= 0.0:
            raise ValueError("LBM_SOLITARY_AMPLITUDE_M must be positive")
        solitary_height_depth_ratio = solitary_amp_m / free_surface_depth_m
    solitary_x0_m = float(env.get("LBM_SOLITARY_X0_M", 0.25 * Lx_m))
    solitary_direction = float(env.get("LBM_SOLITARY_DIRECTION", 1.0))
    if solitary_direction == 0.0:
        raise ValueError("LBM_SOLITARY_DIRECTION must be nonzero")
    solitary_direction = 1.0 if solitary_direction > 0.0 else -1.0
    solitary_initial_velocity_mode = str(env.get("LBM_SOLITARY_INITIAL_VELOCITY", "shallow_water_divfree")).lower()
    if solitary_initial_velocity_mode not in ("none", "shallow_water", "shallow_water_divfree"):
        raise ValueError(
            "LBM_SOLITARY_INITIAL_VELOCITY must be 'none', 'shallow_water', "
            "or 'shallow_water_divfree'"
        )
    solitary_celerity_phys = (
        float(np.sqrt(gravity_phys * (free_surface_depth_m + solitary_amp_m))) if gravity_phys > 0.0 else 0.0
    )
    solitary_kappa_m_inv = float(np.sqrt(3.0 * solitary_amp_m / (4.0 * free_surface_depth_m**3)))
    gaussian_amp_m = float(env.get("LBM_GAUSSIAN_AMPLITUDE_M", 3.0))
    gaussian_center_x_m = float(env.get("LBM_GAUSSIAN_CENTER_X_M", 0.5 * Lx_m))
    gaussian_sigma_m = float(env.get("LBM_GAUSSIAN_SIGMA_M", 12.5))
    gaussian_celerity_phys = (
        float(np.sqrt(gravity_phys * (free_surface_depth_m + gaussian_amp_m))) if gravity_phys > 0.0 else 0.0
    )
    vof_block_size_m = float(env.get("LBM_VOF_BLOCK_SIZE_M", 4.0))
    vof_block_size_x_m = float(env.get("LBM_VOF_BLOCK_SIZE_X_M", vof_block_size_m))
    vof_block_size_y_m = float(env.get("LBM_VOF_BLOCK_SIZE_Y_M", min(vof_block_size_m, Ly_m)))
    vof_block_size_z_m = float(env.get("LBM_VOF_BLOCK_SIZE_Z_M", vof_block_size_m))
    vof_block_center_x_m = float(env.get("LBM_VOF_BLOCK_CENTER_X_M", 0.5 * Lx_m))
    vof_block_center_y_m = float(env.get("LBM_VOF_BLOCK_CENTER_Y_M", 0.5 * Ly_m))
    vof_block_bottom_default = min(Lz_m - vof_block_size_z_m - dx, free_surface_base_m + 2.0 * dx)
    vof_block_bottom_m = float(env.get("LBM_VOF_BLOCK_BOTTOM_M", vof_block_bottom_default))
    vof_block_center_z_m = float(env.get("LBM_VOF_BLOCK_CENTER_Z_M", vof_block_bottom_m + 0.5 * vof_block_size_z_m))
    if vof_initial_shape != "none" and tracking != "vof":
        raise ValueError("LBM_VOF_INITIAL_SHAPE requires LBM_FREE_SURFACE_TRACKING=vof")
    vof_disconnected = vof_initial_shape != "none"
    allow_total = _flag(env, "LBM_ALLOW_TOTAL_PRESSURE_VOF_DIAGNOSTIC", "0")
    if vof_disconnected and pressure_formulation == "total_pressure" and not allow_total:
        pressure_formulation = "vof_component_balanced"
        pf_auto = True
    if vof_disconnected and pressure_formulation == "hydrostatic_balanced":
        raise ValueError(
            "Disconnected VOF features with LBM_PRESSURE_FORMULATION=hydrostatic_balanced are not supported; "
            "use vof_component_balanced so the flat pool is balanced while detached water still falls."
        )
    if vof_disconnected and fs_boundary == "hydrostatic":
        raise ValueError(
            "Disconnected VOF features require LBM_FREE_SURFACE_BOUNDARY=cell; "
            "the hydrostatic option still uses a column-height surface."
        )
    vof_block_drop_height_m = max(vof_block_center_z_m - free_surface_base_m, 0.0)
    vof_reference_speed_phys = (
        float(np.sqrt(2.0 * gravity_phys * vof_block_drop_height_m))
        if (gravity_phys > 0.0 and vof_initial_shape != "none")
        else 0.0
    )
    obstacle_mode = str(env.get("LBM_OBSTACLE_MODE", "bed")).lower()
    if obstacle_mode not in ("bed", "bed_cube", "cube"):
        raise ValueError("LBM_OBSTACLE_MODE must be 'bed', 'bed_cube', or 'cube'")
    include_bed_geometry = 1 if obstacle_mode in ("bed", "bed_cube") else 0
    include_cube_geometry = 1 if obstacle_mode in ("bed_cube", "cube") else 0
    # the box obstacle (CODEX:376-400): sizes, centre, yaw about z, pitch about y
    cube_side_m = float(env.get("LBM_CUBE_SIDE_M", 6.0))
    cube_size = [float(env.get(f"LBM_CUBE_SIZE_{a}_M", cube_side_m)) for a in "XYZ"]
    if min(cube_size) <= 0.0:
        raise ValueError("LBM_CUBE_SIZE_X_M, LBM_CUBE_SIZE_Y_M, and LBM_CUBE_SIZE_Z_M must be positive")
    cube_half = [0.5 * v for v in cube_size]
    cube_yaw_deg = float(env.get("LBM_CUBE_YAW_DEG", 0.0))
    cube_pitch_deg = float(env.get("LBM_CUBE_PITCH_DEG", 0.0))
    cube_rot = [float(np.cos(np.deg2rad(cube_yaw_deg))), float(np.sin(np.deg2rad(cube_yaw_deg))),
                float(np.cos(np.deg2rad(cube_pitch_deg))), float(np.sin(np.deg2rad(cube_pitch_deg)))]
    cube_vertical_half_extent_m = cube_half[2] * abs(cube_rot[2]) + cube_half[0] * abs(cube_rot[3])
    cube_center_z_default = min(
        Lz_m - cube_vertical_half_extent_m - dx,
        max(0.5 * Lz_m, offset_m + 2.0 * bed_amp_m + cube_vertical_half_extent_m + dx),
    )
    cube_center = [float(env.get("LBM_CUBE_CENTER_X_M", 0.5 * Lx_m)), float(env.get("LBM_CUBE_CENTER_Y_M", 0.5 * Ly_m)),
                   float(env.get("LBM_CUBE_CENTER_Z_M", cube_center_z_default))]

    rho0 = 1.0
    u_lid = 0.075
    U_ref_phys = max(
        abs(U_top_phys),
        abs(U_wave_phys),
        abs(water_current_phys),
        abs(solitary_celerity_phys),
        abs(gaussian_celerity_phys),
        abs(vof_reference_speed_phys),
        1e-6,
    )
    vel_scale = U_ref_phys / u_lid
    dt_phys = u_lid * dx / U_ref_phys
    u_top_lb = U_top_phys / vel_scale
    u_bed_lb = wave_speed / vel_scale
    u_current_lb = water_current_phys / vel_scale
    g_lb = gravity_phys * dt_phys * dt_phys / dx
    if "LBM_G_LB" in env:
        g_lb = float(env["LBM_G_LB"])
    use_cbp = 1 if (water_free_surface_enabled == 1 and pressure_formulation == "vof_component_balanced") else 0
    use_hbp = (
        1
        if (
            water_free_surface_enabled == 1
            and pressure_formulation in ("hydrostatic_balanced", "vof_component_balanced")
        )
        else 0
    )
    g_body_lb = 0.0 if (water_free_surface_enabled == 1 and pressure_formulation == "hydrostatic_balanced") else g_lb
    Re_phys_air = U_ref_phys * Lz_m / nu_phys
    nu0 = u_lid * Nz / Re_phys_air
    tau0 = 0.5 + 3 * nu0
    if "LBM_TAU" in env:
        tau0 = float(env["LBM_TAU"])
        if not tau0 > 0.5:
            raise ValueError("LBM_TAU must be greater than 0.5")
        nu0 = (tau0 - 0.5) / 3.0
    Re_lb_base = u_lid * Nz / nu0 if nu0 > 0 else float("inf")

    use_LES = _flag(env, "LBM_USE_LES", "1" if physics_mode == "air" else "0")
    collision_model = str(env.get("LBM_COLLISION", "regularized" if physics_mode == "water" else "kbc")).lower()
    if collision_model not in ("kbc", "regularized"):
        raise ValueError("LBM_COLLISION must be 'kbc' or 'regularized'")
    Cs = float(env.get("LBM_LES_CS", 0.2))
    Delta = 1.0
    nu_t_max = 10.0
    TI = 0.1
    kappa = 0.41
    z0_phys = 0.01
    C_backscatter = float(env.get("LBM_BACKSCATTER", 0.02 if physics_mode == "air" else 0.0))
    enable_top_drive = 1 if (physics_mode == "air" and env.get("LBM_ENABLE_TOP_DRIVE", "1") != "0") else 0
    enable_x_fringe = (
        1
        if (physics_mode == "air" and x_boundary_mode == "periodic" and env.get("LBM_ENABLE_X_FRINGE", "1") != "0")
        else 0
    )
    enable_top_sponge = 1 if (physics_mode == "air" and env.get("LBM_ENABLE_TOP_SPONGE", "1") != "0") else 0

    top_boundary_mode = env_choice(env, "LBM_TOP_BOUNDARY", "keep", ("keep", "free_slip"))
    bed_phase_speed_lu = float(env.get("LBM_BED_PHASE_SPEED_LU", u_bed_lb))
    bed_phase_speed_mps = bed_phase_speed_lu * vel_scale
    storage = env_choice(env, "LBM_STORAGE", "fp32", ("fp32", "fp16", "bf16", "fp64"))
    storage_rounding = env_choice(env, "LBM_STORAGE_ROUNDING", "stochastic", ("stochastic", "nearest"))
    kernel_path = env_choice(env, "LBM_KERNEL_PATH", "auto", ("auto", "staged", "fused"))
    initial_state = env_choice(
        env, "LBM_INITIAL_STATE", "reference", ("reference", "uniform_noise", "loglaw_noise")
    )
    init_u_lu = float(env.get("LBM_INIT_U_LU", 0.05))
    init_noise_lu = float(env.get("LBM_INIT_NOISE_LU", 1e-3))
    inflow_u_lu = float(env.get("LBM_INFLOW_U_LU", 0.06))
    inflow_z0_cells = float(env.get("LBM_INFLOW_Z0_CELLS", 0.5))
    inflow_noise = float(env.get("LBM_INFLOW_NOISE", 0.02))
    seed = env_int(env, "LBM_SEED", 12345)

    u_ref_text = str(env.get("LBM_STORAGE_U_REF", "auto")).lower()
    if u_ref_text == "auto":
        storage_u_ref = auto_storage_u_ref(physics_mode, initial_state, init_u_lu, inflow_u_lu, u_top_lb)
    else:
        parts = [float(x) for x in u_ref_text.split(",")]
        if len(parts) == 1:
            parts = [parts[0], 0.0, 0.0]
        if len(parts) != 3:
            raise ValueError("LBM_STORAGE_U_REF must be 'auto', one value or 'ux,uy,uz'")
        storage_u_ref = tuple(parts)
    if max(abs(x) for x in storage_u_ref) > 0.3:
        raise ValueError("LBM_STORAGE_U_REF components must lie in [-0.3, 0.3]")
    storage_shift = env_int(env, "LBM_STORAGE_SHIFT", -1)
    if storage_shift < 0:
        speeds = [abs(init_u_lu) + init_noise_lu, abs(inflow_u_lu) * (1.0 + inflow_noise), abs(u_top_lb)]
        storage_shift = default_storage_shift(storage, max(speeds), storage_u_ref)
    if storage_shift > 24:
        raise ValueError("LBM_STORAGE_SHIFT must be in [0, 24]")

    # features this build does not implement are refused, not ignored
    unsupported = []

    v.update(
        dict(
            physics_mode=physics_mode,
            free_surface_mode=free_surface_mode,
            water_free_surface_enabled=water_free_surface_enabled,
            free_surface_initial_condition=fs_initial,
            vof_initial_shape=vof_initial_shape,
            initial_pressure_mode=initial_pressure_mode,
            pressure_formulation=pressure_formulation,
            pressure_formulation_auto_corrected=pf_auto,
            free_surface_tracking_mode=tracking,
            free_surface_boundary_mode=fs_boundary,
            free_surface_refill_mode=refill,
            initialize_free_surface_populations=init_fs_pops,
            vof_link_aperture_enabled=vof_link_aperture_enabled,
            vof_empty_fill_threshold=vof_empty,
            vof_orphan_empty_threshold=vof_orphan_thr,
            vof_orphan_search_radius=vof_orphan_r,
            vof_orphan_max_weak_neighbors=vof_orphan_weak,
            vof_detached_advection_enabled=vof_det_adv,
            vof_detached_max_wet_neighbors=vof_det_wet,
            vof_detached_max_resolved_neighbors=vof_det_res,
            vof_detached_residual_fill_threshold=vof_det_resid,
            vof_thin_gap_bridge_enabled=vof_bridge,
            vof_thin_gap_bridge_strength=vof_bridge_strength,
            vof_collapse_trapped_gas_enabled=vof_collapse,
            vof_collapse_interval=vof_collapse_interval,
            vof_collapse_max_volume_m3=vof_collapse_max_volume_m3,
            vof_collapse_flood_sweeps=vof_collapse_flood_sweeps,
            U_top_phys=U_top_phys,
            water_current_phys=water_current_phys,
            Lx_m=Lx_m,
            Ly_m=Ly_m,
            Lz_m=Lz_m,
            dx=dx,
            nu_phys=nu_phys,
            rho_phys=rho_phys,
            rho_air_phys=rho_phys,
            fluid_name=fluid_name,
            gravity_phys=gravity_phys,
            Nx=Nx,
            Ny=Ny,
            Nz=Nz,
            steps=steps,
            plot_freq=plot_freq,
            plot_step_start=plot_step_start,
            boundary_geometry_mode=boundary_geometry_mode,
            x_boundary_mode=x_boundary_mode,
            y_boundary_mode=y_boundary_mode,
            top_boundary_mode=top_boundary_mode,
            wall_velocity_sampling=wall_velocity_sampling,
            use_link_wall_velocity=use_link_wall_velocity,
            bed_profile=bed_profile,
            bed_is_moving=bed_is_moving,
            bed_amp_m=bed_amp_m,
            depth_swl_m=depth_swl_m,
            wave_speed=wave_speed,
            bed_wavelength_m=bed_wavelength_m,
            wave_period=wave_period,
            offset_m=offset_m,
            U_wave_phys=U_wave_phys,
            piecewise_bed_z_np=piecewise_bed_z_np,
            piecewise_bed_metadata=piecewise_meta,
            free_surface_depth_m=free_surface_depth_m,
            free_surface_base_m=free_surface_base_m,
            solitary_amp_m=solitary_amp_m,
            solitary_height_depth_ratio=solitary_height_depth_ratio,
            solitary_x0_m=solitary_x0_m,
            solitary_direction=solitary_direction,
            solitary_initial_velocity_mode=solitary_initial_velocity_mode,
            solitary_celerity_phys=solitary_celerity_phys,
            solitary_kappa_m_inv=solitary_kappa_m_inv,
            solitary_length_scale_m=1.0 / solitary_kappa_m_inv,
            gaussian_amp_m=gaussian_amp_m,
            gaussian_center_x_m=gaussian_center_x_m,
            gaussian_sigma_m=gaussian_sigma_m,
            vof_block_size_x_m=vof_block_size_x_m,
            vof_block_size_y_m=vof_block_size_y_m,
            vof_block_size_z_m=vof_block_size_z_m,
            vof_block_center_x_m=vof_block_center_x_m,
            vof_block_center_y_m=vof_block_center_y_m,
            vof_block_center_z_m=vof_block_center_z_m,
            obstacle_mode=obstacle_mode,
            include_bed_geometry=include_bed_geometry,
            include_cube_geometry=include_cube_geometry,
            cube_size_x_m=cube_size[0], cube_size_y_m=cube_size[1], cube_size_z_m=cube_size[2], cube_side_m=cube_side_m,
            cube_half_x_m=cube_half[0], cube_half_y_m=cube_half[1], cube_half_z_m=cube_half[2], cube_half=cube_half,
            cube_center_x_m=cube_center[0], cube_center_y_m=cube_center[1], cube_center_z_m=cube_center[2],
            cube_center=cube_center, cube_yaw_deg=cube_yaw_deg, cube_pitch_deg=cube_pitch_deg,
            cube_yaw_cos=cube_rot[0], cube_yaw_sin=cube_rot[1], cube_pitch_cos=cube_rot[2], cube_pitch_sin=cube_rot[3],
            cube_rot=cube_rot,
            rho0=rho0,
            u_lid=u_lid,
            U_ref_phys=U_ref_phys,
            vel_scale=vel_scale,
            dt_phys=dt_phys,
            u_top_lb=u_top_lb,
            u_bed_lb=u_bed_lb,
            u_current_lb=u_current_lb,
            g_lb=g_lb,
            g_body_lb=g_body_lb,
            use_component_balanced_pressure=use_cbp,
            use_hydrostatic_balanced_pressure=use_hbp,
            Re_phys_air=Re_phys_air,
            Re_phys=Re_phys_air,
            nu0=nu0,
            tau0=tau0,
            Re_lb_base=Re_lb_base,
            use_LES=use_LES,
            collision_model=collision_model,
            Cs=Cs,
            Delta=Delta,
            nu_t_max=nu_t_max,
            TI=TI,
            kappa=kappa,
            z0_phys=z0_phys,
            C_backscatter=C_backscatter,
            enable_top_drive=enable_top_drive,
            enable_x_fringe=enable_x_fringe,
            enable_top_sponge=enable_top_sponge,
            bed_phase_speed_lu=bed_phase_speed_lu,
            bed_phase_speed_mps=bed_phase_speed_mps,
            storage=storage,
            storage_shift=storage_shift,
            storage_rounding=storage_rounding,
            storage_u_ref=storage_u_ref,
            kernel_path=kernel_path,
            initial_state=initial_state,
            init_u_lu=init_u_lu,
            init_noise_lu=init_noise_lu,
            inflow_u_lu=inflow_u_lu,
            inflow_z0_cells=inflow_z0_cells,
            inflow_noise=inflow_noise,
            seed=seed,
            unsupported=unsupported,
        )
    )
    return SolverConfig(values=v)


def auto_storage_u_ref(physics_mode, initial_state, init_u_lu, inflow_u_lu, u_top_lb):
    """Reference velocity of the reduced-storage shift: the run's mean flow
    (x direction), so that the stored deviations f_i - feq_i(1, u_ref) of
    the bulk are small and keep their relative precision."""
    if physics_mode == "water":
        return (0.0, 0.0, 0.0)
    if initial_state == "uniform_noise":
        return (float(init_u_lu), 0.0, 0.0)
    if initial_state == "loglaw_noise":
        return (float(inflow_u_lu), 0.0, 0.0)
    return (float(np.float32(0.8) * np.float32(u_top_lb)), 0.0, 0.0)


def default_storage_shift(storage: str, speed: float = 0.1, u_ref=(0.0, 0.0, 0.0)) -> int:
    """k of the (f_i - s_i) * 2^k scaling, chosen from a bound of the largest
    deviation.  With |u - u_ref| <= U = 2 * (speed + |u_ref|) and density
    excursions of 5 %, |f_i - s_i| <= D = (3U + 6U^2) / 18 + 0.05 / 3; k puts
    D at 2^12 so the stored values keep a 16x margin below the largest half
    (65504) while deviations down to D * 2^-26 stay in half's normal range.
    bfloat16 has float32's exponent range and is not scaled."""
    if storage != "fp16":
        return 0
    U = 2.0 * (abs(speed) + math.sqrt(sum(x * x for x in u_ref)))
    D = (3.0 * U + 6.0 * U * U) / 18.0 + 0.05 / 3.0
    return int(min(24, max(0, math.floor(math.log2(4096.0 / D)))))


def bed_max_height(cfg: SolverConfig) -> float:
    """Upper bound of the bed height over all time (used to bound the band of
    cells that can change solid/fluid state)."""
    if cfg.bed_profile == "moving_sine":
        return cfg.offset_m + 2.0 * abs(cfg.bed_amp_m)
    if cfg.bed_profile == "moving_sine1":
        return cfg.offset_m + abs(cfg.bed_amp_m)
    if cfg.bed_profile == "piecewise_linear":
        return float(np.max(cfg.piecewise_bed_z_np))
    return cfg.offset_m


# the benchmark workloads as environment presets ----------------------------
def preset_env(name: str) -> dict:
    """LBM_* settings of the benchmark workloads (all lattice-unit choices are
    made with dx = 1 m so that cells and metres coincide)."""
    if name == "plain_64x32x32":
        return {
            "LBM_PHYSICS_MODE": "air",
            "LBM_COLLISION": "regularized",
            "LBM_USE_LES": "1",
            "LBM_LES_CS": "0.16",
            "LBM_TAU": "0.52",
            "LBM_DX_M": "1",
            "LBM_LX_M": "64",
            "LBM_LY_M": "32",
            "LBM_LZ_M": "32",
            "LBM_BED_PROFILE": "flat",
            "LBM_BED_LEVEL_M": "1.25",
            "LBM_X_BOUNDARY": "periodic",
            "LBM_TOP_BOUNDARY": "free_slip",
            "LBM_ENABLE_TOP_DRIVE": "0",
            "LBM_ENABLE_X_FRINGE": "0",
            "LBM_ENABLE_TOP_SPONGE": "0",
            "LBM_BACKSCATTER": "0",
            "LBM_INITIAL_STATE": "uniform_noise",
            "LBM_INIT_U_LU": "0.05",
            "LBM_INIT_NOISE_LU": "1e-3",
            "LBM_SEED": "12345",
        }
    if name == "waves_256x128x128":
        return {
            "LBM_PHYSICS_MODE": "air",
            "LBM_COLLISION": "kbc",
            "LBM_USE_LES": "1",
            "LBM_LES_CS": "0.16",
            "LBM_TAU": "0.52",
            "LBM_DX_M": "1",
            "LBM_LX_M": "256",
            "LBM_LY_M": "128",
            "LBM_LZ_M": "128",
            "LBM_BED_PROFILE": "moving_sine1",
            "LBM_BED_AMP_M": "4",
            "LBM_BED_WAVELENGTH_M": "64",
            "LBM_BED_PHASE_SPEED_LU": "0.02",
            "LBM_X_BOUNDARY": "inflow_outflow",
            "LBM_TOP_BOUNDARY": "free_slip",
            "LBM_ENABLE_TOP_DRIVE": "0",
            "LBM_ENABLE_TOP_SPONGE": "0",
            "LBM_BACKSCATTER": "0",
            "LBM_INITIAL_STATE": "loglaw_noise",
            "LBM_INFLOW_U_LU": "0.06",
            "LBM_INFLOW_Z0_CELLS": "0.5",
            "LBM_INFLOW_NOISE": "0.02",
            "LBM_SEED": "12345",
        }
    if name == "plain_512x256x192":
        env = preset_env("plain_64x32x32")
        env.update({"LBM_LX_M": "512", "LBM_LY_M": "256", "LBM_LZ_M": "192", "LBM_STORAGE": "fp16"})
        return env
    if name == "vof_192x96x160":
        return {
            "LBM_PHYSICS_MODE": "water",
            "LBM_FREE_SURFACE_MODE": "height_kinematic",
            "LBM_FREE_SURFACE_TRACKING": "vof",
            "LBM_FREE_SURFACE_INITIAL": "flat",
            "LBM_PRESSURE_FORMULATION": "vof_component_balanced",
            "LBM_COLLISION": "regularized",
            "LBM_USE_LES": "1",
            "LBM_LES_CS": "0.16",
            "LBM_TAU": "0.52",
            "LBM_DX_M": "1",
            "LBM_LX_M": "192",
            "LBM_LY_M": "96",
            "LBM_LZ_M": "160",
            "LBM_BED_PROFILE": "flat",
            "LBM_BED_LEVEL_M": "0",
            "LBM_FREE_SURFACE_DEPTH_M": "80",
            "LBM_X_BOUNDARY": "solid",
            "LBM_Y_BOUNDARY": "solid",
            "LBM_VOF_INITIAL_SHAPE": "block",
            "LBM_VOF_BLOCK_SIZE_M": "16",
            "LBM_VOF_BLOCK_BOTTOM_M": "120",
            "LBM_G_LB": "1e-5",
            "LBM_VOF_COLLAPSE_TRAPPED_GAS": "1",
            "LBM_VOF_COLLAPSE_INTERVAL": "200",
            "LBM_VOF_COLLAPSE_MAX_VOLUME_M3": "512",
            "LBM_SEED": "12345",
        }
    raise KeyError(name)
