"""ctypes mirror of include/lbm.h (config struct, counters, constants).

Shared by the solver facade (lbm_solver.py), the ensemble server and the
tests.  The struct layouts must match include/lbm.h field for field;
tests/test_abi_config.py checks sizes and offsets against the C header.
"""
from __future__ import annotations

import ctypes

LBM_CONFIG_VERSION = 8

# error codes
LBM_OK = 0
LBM_E_ARG = -1
LBM_E_CONFIG = -2
LBM_E_HIP = -3
LBM_E_NOMEM = -4
LBM_E_STATE = -5
LBM_E_UNSUPPORTED = -6

PHYS_AIR, PHYS_WATER = 0, 1
COLL_KBC, COLL_REGULARIZED = 0, 1
XB_PERIODIC, XB_SOLID, XB_INFLOW_OUTFLOW = 0, 1, 2
YB_PERIODIC, YB_SOLID = 0, 1
TOP_KEEP, TOP_FREE_SLIP = 0, 1
BED_MOVING_SINE, BED_FLAT, BED_PIECEWISE, BED_MOVING_SINE1 = 0, 1, 2, 3
STORE_FP32, STORE_FP16, STORE_BF16, STORE_FP64 = 0, 1, 2, 3
ROUND_NEAREST, ROUND_STOCHASTIC = 0, 1
PATH_AUTO, PATH_STAGED, PATH_FUSED = 0, 1, 2
FSM_NONE, FSM_HEIGHT_STATIC, FSM_HEIGHT_KINEMATIC, FSM_PRESCRIBED_SOLITARY = 0, 1, 2, 3
FST_HEIGHT, FST_VOF = 0, 1
PF_TOTAL, PF_HYDROSTATIC_BALANCED, PF_VOF_COMPONENT_BALANCED = 0, 1, 2
FSB_CELL, FSB_HYDROSTATIC = 0, 1
REFILL_EQUILIBRIUM, REFILL_NONEQ, REFILL_DIRECTIONAL = 0, 1, 2

WALL_NONE, WALL_BED, WALL_CUBE, WALL_XMIN, WALL_XMAX, WALL_YMIN, WALL_YMAX = 0, 1, 2, 3, 4, 5, 6
FS_GAS, FS_LIQUID, FS_INTERFACE = 0, 1, 2

# field identifiers (include/lbm.h)
FIELD_F = 0
FIELD_RHO = 1
FIELD_UX = 2
FIELD_UY = 3
FIELD_UZ = 4
FIELD_OMEGA = 5
FIELD_PHI = 6
FIELD_LATTICE_OPEN = 7
FIELD_LATTICE_OPEN_FRAC = 8
FIELD_WALL_TYPE = 9
FIELD_NEAR_OBSTACLE = 10
FIELD_OBSTACLE = 11
FIELD_WATER_H = 12
FIELD_U_WAVE = 13
FIELD_V_WAVE = 14
FIELD_W_WAVE = 15
FIELD_FS_FILL = 16
FIELD_FS_TYPE = 17
FIELD_FS_TYPE_PREV = 18
FIELD_FS_MASS = 19
FIELD_FS_MASS_EXCESS = 20
FIELD_FS_H = 21
FIELD_PRESSURE_REF_HEAD = 22
FIELD_GRAVITY_FRACTION = 23
FIELD_U_INLET_PROFILE = 24
FIELD_INFLOW_UX = 25
FIELD_PIECEWISE_BED_Z = 26
FIELD_NU_T = 27
FIELD_SMAG = 28
FIELD_VOF_EXTERIOR_GAS = 29
FIELD_COUNT = 30

INT_FIELDS = {
    FIELD_LATTICE_OPEN,
    FIELD_WALL_TYPE,
    FIELD_NEAR_OBSTACLE,
    FIELD_OBSTACLE,
    FIELD_FS_TYPE,
    FIELD_FS_TYPE_PREV,
    FIELD_VOF_EXTERIOR_GAS,
}
LINK_FIELDS = {FIELD_F, FIELD_LATTICE_OPEN, FIELD_LATTICE_OPEN_FRAC, FIELD_WALL_TYPE}
PLANE_FIELDS = {FIELD_WATER_H, FIELD_U_WAVE, FIELD_V_WAVE, FIELD_W_WAVE, FIELD_FS_H}

REDUCE_MASS = 0
REDUCE_VOF_MASS = 1
REDUCE_VOF_VOLUME = 2
REDUCE_MAX_SPEED = 3
REDUCE_NONFINITE = 4


def field_shape(field: int, nx: int, ny: int, nz: int) -> tuple:
    """Host array shape of a field, C order with x fastest."""
    if field in LINK_FIELDS:
        return (19, nz, ny, nx)
    if field in PLANE_FIELDS:
        return (ny, nx)
    if field == FIELD_U_INLET_PROFILE:
        return (nz,)
    if field == FIELD_INFLOW_UX:
        return (nz, ny)
    if field == FIELD_PIECEWISE_BED_Z:
        return (nx,)
    if 0 <= field < FIELD_COUNT:
        return (nz, ny, nx)
    raise ValueError(f"unknown field id {field}")


_I = ctypes.c_int32
_F = ctypes.c_float


class LbmConfig(ctypes.Structure):
    _fields_ = [
        ("version", _I),
        ("nx", _I),
        ("ny", _I),
        ("nz", _I),
        ("physics_mode", _I),
        ("collision", _I),
        ("use_les", _I),
        ("x_boundary", _I),
        ("y_boundary", _I),
        ("top_boundary", _I),
        ("bed_profile", _I),
        ("include_bed", _I),
        ("wall_velocity_link", _I),
        ("enable_top_drive", _I),
        ("enable_x_fringe", _I),
        ("enable_top_sponge", _I),
        ("storage", _I),
        ("storage_shift", _I),
        ("storage_rounding", _I),
        ("kernel_path", _I),
        ("fs_enabled", _I),
        ("fs_tracking", _I),
        ("pressure_formulation", _I),
        ("fs_boundary_mode", _I),
        ("refill_mode", _I),
        ("vof_link_aperture", _I),
        ("vof_thin_gap_bridge", _I),
        ("vof_detached_advection", _I),
        ("vof_collapse_trapped_gas", _I),
        ("vof_orphan_search_radius", _I),
        ("vof_orphan_max_weak_neighbors", _I),
        ("vof_detached_max_wet_neighbors", _I),
        ("vof_detached_max_resolved_neighbors", _I),
        ("vof_collapse_interval", _I),
        ("vof_collapse_flood_sweeps", _I),
        ("fs_mode", _I),
        ("include_cube", _I),
        ("phi_fraction", _I),
        ("phi_flags", _I),
        ("init_fs_populations", _I),
        ("seed", ctypes.c_uint32),
        ("dx", _F),
        ("rho0", _F),
        ("tau0", _F),
        ("nu0", _F),
        ("cs_les", _F),
        ("delta_les", _F),
        ("nu_t_max", _F),
        ("kappa", _F),
        ("z0_phys", _F),
        ("u_top_lb", _F),
        ("u_current_lb", _F),
        ("g_lb", _F),
        ("g_body_lb", _F),
        ("c_backscatter", _F),
        ("offset_m", _F),
        ("bed_amp_m", _F),
        ("bed_wavelength_m", _F),
        ("wave_speed", _F),
        ("wave_period", _F),
        ("depth_swl_m", _F),
        ("vel_scale", _F),
        ("bed_phase_speed_mps", _F),
        ("free_surface_base_m", _F),
        ("vof_empty_fill_threshold", _F),
        ("vof_orphan_empty_threshold", _F),
        ("vof_detached_residual_fill_threshold", _F),
        ("vof_thin_gap_bridge_strength", _F),
        ("vof_collapse_max_volume_m3", _F),
        ("storage_u_ref", _F * 3),
        ("cube_half", _F * 3),
        ("cube_center", _F * 3),
        ("cube_rot", _F * 4),
        ("solitary", _F * 5),
        ("pad0", _F),
        ("dt_phys", ctypes.c_double),
    ]

    def copy(self) -> "LbmConfig":
        out = LbmConfig()
        ctypes.memmove(ctypes.byref(out), ctypes.byref(self), ctypes.sizeof(LbmConfig))
        return out


LBM_INIT_PARAMS_VERSION = 1
INIT_REFERENCE, INIT_UNIFORM_NOISE, INIT_LOGLAW_NOISE = 0, 1, 2
FSI_FLAT, FSI_GAUSSIAN, FSI_SOLITARY = 0, 1, 2
SOLVEL_NONE, SOLVEL_SHALLOW, SOLVEL_SHALLOW_DIVFREE = 0, 1, 2

_D = ctypes.c_double


class LbmInitParams(ctypes.Structure):
    """lbm_init_params_t: the options of the host half of the init
    sequence (CODEX:3777-3792)."""
    _fields_ = [
        ("version", _I),
        ("initial_state", _I),
        ("fs_initial", _I),
        ("vof_block", _I),
        ("linear_wave_pressure", _I),
        ("solitary_velocity", _I),
        ("hydrostatic_balanced", _I),
        ("component_balanced", _I),
        ("pad0", _I),
        ("seed", ctypes.c_uint64),
        *[(n, _D) for n in (
            "dx", "offset_m", "free_surface_base_m", "free_surface_depth_m", "rho0", "g_lb", "u_current_lb",
            "u_top_lb", "tau0", "kappa", "vel_scale", "turbulence_intensity", "init_u_lu", "init_noise_lu",
            "inflow_u_lu", "inflow_z0_cells", "inflow_noise", "gaussian_amp_m", "gaussian_sigma_m",
            "gaussian_center_x_m", "solitary_amp_m", "solitary_x0_m", "solitary_direction",
            "solitary_celerity_phys", "solitary_kappa_m_inv")],
        ("vof_block_size_m", _D * 3),
        ("vof_block_center_m", _D * 3),
    ]


class LbmCounters(ctypes.Structure):
    _fields_ = [
        ("boundary_cut_link_count", _I),
        ("boundary_refill_count", _I),
        ("free_surface_refill_count", _I),
        ("free_surface_coalescence_repair_count", _I),
        ("free_surface_thin_gap_bridge_count", _I),
        ("free_surface_cut_link_count", _I),
        ("free_surface_vof_excess_count", _I),
        ("free_surface_vof_orphan_count", _I),
        ("free_surface_vof_collapse_cell_count", _I),
        ("free_surface_vof_detached_advect_count", _I),
        ("free_surface_vof_detached_compress_count", _I),
        ("pad0", _I),
        ("free_surface_vof_collapse_candidate_volume_m3", ctypes.c_double),
        ("free_surface_vof_collapse_applied_volume_m3", ctypes.c_double),
        ("free_surface_vof_collapse_removal_scale", ctypes.c_double),
        ("free_surface_vof_collapse_candidate_mass", ctypes.c_double),
        ("free_surface_vof_collapse_exterior_capacity", ctypes.c_double),
        ("free_surface_vof_detached_advect_mass", ctypes.c_double),
        ("free_surface_vof_detached_compress_mass", ctypes.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_ if name != "pad0"}
