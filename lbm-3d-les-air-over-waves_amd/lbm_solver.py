"""Solver facade: the reference solver module's surface over liblbm.so.

The reference (LBM_3D_LES_air_over_waves_CODEX.py, "CODEX") is a module
whose import parses the LBM_* environment, allocates Taichi fields and
defines kernels; main() runs the init sequence (CODEX:3777-3792) and the
step loop (CODEX:3873-3955).  The harness (run_ab_case.py:261-906) reads
module attributes: constants, fields through ``.to_numpy()`` and counters
through ``[None]``, and replaces ``visualize``.

``Solver`` reproduces that surface: the same attribute names, field proxies
with ``to_numpy()`` / ``from_numpy()``, counter proxies with ``[None]``,
``initialize()`` (the init sequence), ``step(n)`` (n loop bodies) and
``main()``.  The module LBM_3D_LES_air_over_waves_hip.py builds one Solver
at import and exposes it as module attributes, so it loads the way the
reference loads.  All compute runs in liblbm.so; there is no CPU path.
"""
from __future__ import annotations

import os
from typing import Callable, Mapping, Optional

import numpy as np

import lbm_abi as abi
from lbm_config import SolverConfig, parse_env
from lbm_init import initial_fields
from lbm_native import Case

D3Q19_CX = (0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0)
D3Q19_CY = (0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, -1, 1)
D3Q19_CZ = (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1)


class ConstField:
    """Read-only stand-in for a constant Taichi field (cx, cy, cz, w, opp)."""

    def __init__(self, values, dtype):
        self._a = np.asarray(values, dtype=dtype)
        self.shape = self._a.shape

    def to_numpy(self):
        return self._a.copy()

    def __getitem__(self, idx):
        return self._a[idx]


class FieldProxy:
    """A device field with the Taichi field methods the harness uses."""

    def __init__(self, solver: "Solver", field: int):
        self._s = solver
        self._field = field
        c = solver.case.cfg
        self.shape = abi.field_shape(field, c.nx, c.ny, c.nz)

    def to_numpy(self) -> np.ndarray:
        a = self._s.case.download(self._field)
        if self._field in abi.LINK_FIELDS:
            # Taichi Vector.field(19) layout: (nz, ny, nx, 19)
            return np.ascontiguousarray(np.moveaxis(a, 0, -1))
        return a

    def from_numpy(self, arr: np.ndarray):
        a = np.asarray(arr)
        if self._field in abi.LINK_FIELDS and a.shape[-1] == 19 and a.ndim == 4:
            a = np.moveaxis(a, -1, 0)
        self._s.case.upload(self._field, a)


class CounterProxy:
    """A shape-() counter read with ``counter[None]``."""

    def __init__(self, solver: "Solver", name: str):
        self._s = solver
        self._name = name

    def __getitem__(self, key):
        if key is not None:
            raise IndexError("scalar counters are indexed with [None]")
        return self._s.counters()[self._name]


_FIELDS = {
    "f": abi.FIELD_F,
    "rho": abi.FIELD_RHO,
    "ux": abi.FIELD_UX,
    "uy": abi.FIELD_UY,
    "uz": abi.FIELD_UZ,
    "omegaLoc": abi.FIELD_OMEGA,
    "phi": abi.FIELD_PHI,
    "lattice_open": abi.FIELD_LATTICE_OPEN,
    "lattice_open_frac": abi.FIELD_LATTICE_OPEN_FRAC,
    "lattice_wall_type": abi.FIELD_WALL_TYPE,
    "near_obstacle": abi.FIELD_NEAR_OBSTACLE,
    "obstacle": abi.FIELD_OBSTACLE,
    "water_h": abi.FIELD_WATER_H,
    "u_wave": abi.FIELD_U_WAVE,
    "v_wave": abi.FIELD_V_WAVE,
    "w_wave": abi.FIELD_W_WAVE,
    "u_inlet_profile": abi.FIELD_U_INLET_PROFILE,
    "piecewise_bed_z": abi.FIELD_PIECEWISE_BED_Z,
    "nu_t_field": abi.FIELD_NU_T,
    "Smag_stress_field": abi.FIELD_SMAG,
    "pressure_reference_head": abi.FIELD_PRESSURE_REF_HEAD,
    "gravity_body_fraction": abi.FIELD_GRAVITY_FRACTION,
}
_WATER_FIELDS = {
    "free_surface_fill": abi.FIELD_FS_FILL,
    "free_surface_type": abi.FIELD_FS_TYPE,
    "free_surface_type_prev": abi.FIELD_FS_TYPE_PREV,
    "free_surface_mass": abi.FIELD_FS_MASS,
    "free_surface_mass_excess": abi.FIELD_FS_MASS_EXCESS,
    "free_surface_h": abi.FIELD_FS_H,
    "vof_exterior_gas": abi.FIELD_VOF_EXTERIOR_GAS,
}
_COUNTERS = [name for name, _ in abi.LbmCounters._fields_ if name != "pad0"]


class Solver:
    """One case: configuration from LBM_* variables, state on the GPU."""

    FS_GAS = abi.FS_GAS
    FS_LIQUID = abi.FS_LIQUID
    FS_INTERFACE = abi.FS_INTERFACE
    WALL_NONE, WALL_BED, WALL_CUBE, WALL_XMIN, WALL_XMAX = 0, 1, 2, 3, 4

    def __init__(self, env: Optional[Mapping[str, str]] = None, config: Optional[SolverConfig] = None):
        self.config = config if config is not None else parse_env(env)
        if self.config.unsupported:
            raise NotImplementedError("not supported by this build: " + ", ".join(self.config.unsupported))
        self.case = Case(self.config.to_struct())
        for name, value in self.config.values.items():
            setattr(self, name, value)
        self.cx = ConstField(D3Q19_CX, np.int32)
        self.cy = ConstField(D3Q19_CY, np.int32)
        self.cz = ConstField(D3Q19_CZ, np.int32)
        dx = self.config.dx
        self.x_phys = (np.arange(self.Nx) + 0.5) * dx
        self.y_phys = (np.arange(self.Ny) + 0.5) * dx
        self.z_phys = (np.arange(self.Nz) + 0.5) * dx
        self.Z_phys, self.Y_phys, self.X_phys = np.meshgrid(self.z_phys, self.y_phys, self.x_phys, indexing="ij")
        for name, fid in _FIELDS.items():
            setattr(self, name, FieldProxy(self, fid))
        if self.config.water_free_surface_enabled:
            for name, fid in _WATER_FIELDS.items():
                setattr(self, name, FieldProxy(self, fid))
        for name in _COUNTERS:
            setattr(self, name, CounterProxy(self, name))
        self.visualize: Callable = self._no_visualize
        self.initialized = False

    # -- the reference's entry points ----------------------------------------
    def initialize(self):
        """The init sequence of main() (CODEX:3777-3792)."""
        cfg = self.config
        case = self.case
        if cfg.bed_profile == "piecewise_linear":
            case.upload(abi.FIELD_PIECEWISE_BED_Z, cfg.piecewise_bed_z_np)
        case.geometry(0.0)
        obstacle0 = case.download(abi.FIELD_OBSTACLE)
        fields = initial_fields(cfg, obstacle0, case.download(abi.FIELD_WATER_H))
        for fid, arr in fields.items():
            case.upload(fid, arr)
        case.init_populations()
        self.initialized = True

    def step(self, n: int = 1):
        """n iterations of the step loop body (CODEX:3874-3925)."""
        if not self.initialized:
            raise RuntimeError("initialize() first")
        self.case.step(n)

    def counters(self) -> dict:
        return self.case.counters()

    @staticmethod
    def _no_visualize(*args, **kwargs):
        raise RuntimeError(
            "Visualization has not been configured. Use run_ab_case.py or run this solver as a script so "
            "direct-run visualization can be bound."
        )

    def health(self) -> float:
        """Max |u| over active cells (the health print, CODEX:3930-3940)."""
        return self.case.reduce(abi.REDUCE_MAX_SPEED)

    def main(self):
        """main() (CODEX:3757-3959): init, step loop, visualize, health print."""
        self.initialize()
        print("Grid Size (Nx, Ny, Nz): ", self.Nx, self.Ny, self.Nz)
        print(f"Physics mode: {self.physics_mode} ({self.fluid_name})")
        print(f"Collision model: {self.collision_model}")
        print(f"Reference speed={self.U_ref_phys:.3f} m/s, dt={self.dt_phys:.6e} s, g_lb={self.g_lb:.6e}")
        print(f"Physical Re ({self.fluid_name}): {self.Re_phys_air:.3e}")
        print(f"Lattice Re (base): {self.Re_lb_base:.2f}")
        step = 0
        while step < self.steps:
            nxt = min(self.steps, step + 100 - step % 100)
            if self.plot_freq > 0:
                nxt = min(nxt, step + self.plot_freq - step % self.plot_freq)
            self.case.step(nxt - step)
            step = nxt
            time_phys = (step + 1) * self.dt_phys
            if self.plot_freq > 0 and step % self.plot_freq == 0 and step > self.plot_step_start:
                self.visualize(step, time_phys, self.Re_phys_air, self.Re_lb_base)
            if step % 100 == 0:
                max_u = self.health()
                cnt = self.counters()
                print(
                    f"Time Step {step}, Max |U| (LB units) = {max_u:.4f}, "
                    f"Max U (phys) ~ {max_u * self.vel_scale:.2f} m/s, "
                    f"cut_links={cnt['boundary_cut_link_count']}, refill={cnt['boundary_refill_count']}, "
                    f"fs_refill={cnt['free_surface_refill_count']}, "
                    f"fs_coalesce={cnt['free_surface_coalescence_repair_count']}, "
                    f"vof_excess={cnt['free_surface_vof_excess_count']}, "
                    f"vof_orphans={cnt['free_surface_vof_orphan_count']}"
                )
        print("Simulation Complete.")

    def close(self):
        self.case.close()


def make_solver(env: Optional[Mapping[str, str]] = None) -> Solver:
    return Solver(env=os.environ if env is None else env)
