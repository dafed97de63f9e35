"""Ensemble server: many independent cases on one GPU, continuously batched.

Requests (a case configuration, a step count and the fields to return)
arrive at any time.  The scheduler admits queued requests first in, first
out while their device footprint (``lbm_device_bytes``) fits the HBM budget
and an active slot is free; it builds the initial state (``Solver.initialize``)
and hands the case to the native stepping runtime (``lbm_server_*`` in
liblbm.so), whose worker threads advance all active cases on their own HIP
streams in chunks of steps, so their device work overlaps.  When the
runtime reports a case finished, the scheduler downloads the requested
fields, frees the case and admits the next requests: a request that arrives
while others run starts as soon as a slot and memory free up.

Every kernel that writes the state is deterministic (no floating-point
atomics on it; FP64 sums in fixed groups and a fixed tree; counter-based
random numbers), so a case's result does not depend on what else runs:
``tests/test_gpu_server.py`` checks alone-vs-mix equality bit for bit,
with a trapped-gas collapse (whose FP64 sums scale the state) in the mix.
"""
from __future__ import annotations

import ctypes
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Mapping, Optional, Sequence

import numpy as np

import lbm_abi as abi
from lbm_config import SolverConfig, parse_env
from lbm_native import LbmError, load_library
from lbm_solver import Solver

DEFAULT_RESULTS = (abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)


@dataclass
class Request:
    ticket: int
    config: SolverConfig
    steps: int
    result_fields: Sequence[int]
    footprint: int
    t_submit: float
    t_admit: float = 0.0
    t_done: float = 0.0
    status: str = "queued"  # queued | active | done | failed
    error: str = ""
    prepare: Optional[Callable] = None  # called with the initialised Solver before stepping
    results: Dict[int, np.ndarray] = field(default_factory=dict)

    @property
    def cells(self) -> int:
        return self.config.Nx * self.config.Ny * self.config.Nz

    @property
    def latency(self) -> float:
        return self.t_done - self.t_submit


def footprint_bytes(config: SolverConfig) -> int:
    lib = load_library()
    out = ctypes.c_size_t()
    cfg = config.to_struct()
    rc = lib.lbm_device_bytes(ctypes.byref(cfg), ctypes.byref(out))
    if rc != 0:
        raise LbmError(f"lbm_device_bytes rejected the configuration ({rc})")
    return int(out.value)


class EnsembleServer:
    """Continuous-batching scheduler over the native stepping runtime."""

    def __init__(self, hbm_budget_bytes: int, max_active: int = 64, chunk_steps: int = 16, n_workers: int = 4,
                 graphs: bool = False, batch: bool = False):
        """graphs: the chunks of the cases on the fused path are captured
        into HIP graphs and launched as one (option "graph" of the C ABI).
        batch: the fused-path cases that share a storage type and collision
        model advance together, one launch per step for all of them
        (lbm_server_batch); the others keep their own streams."""
        self.lib = load_library()
        self.graphs = bool(graphs)
        self.budget = int(hbm_budget_bytes)
        self.max_active = int(max_active)
        self.srv = ctypes.c_void_p()
        rc = self.lib.lbm_server_create(int(chunk_steps), int(n_workers), ctypes.byref(self.srv))
        if rc != 0:
            raise LbmError(f"lbm_server_create failed ({rc})")
        if batch:
            rc = self.lib.lbm_server_batch(self.srv, 1)
            if rc != 0:
                self.lib.lbm_server_destroy(self.srv)
                raise LbmError(f"lbm_server_batch failed ({rc})")
        self.queue: deque = deque()
        self.active: Dict[int, tuple] = {}
        self.finished: List[Request] = []
        self.used = 0
        self.peak_used = 0
        self.peak_active = 0
        self._next = 0

    # -- client side ---------------------------------------------------------
    def submit(self, env_or_config, steps: int, result_fields: Sequence[int] = DEFAULT_RESULTS,
               prepare: Optional[Callable] = None) -> int:
        """Queue a case; `prepare(solver)`, if given, runs after the
        initialisation (e.g. uploads of an initial state of its own)."""
        config = env_or_config if isinstance(env_or_config, SolverConfig) else parse_env(env_or_config)
        if config.unsupported:
            raise NotImplementedError("not supported by this build: " + ", ".join(config.unsupported))
        fp = footprint_bytes(config)
        if fp > self.budget:
            raise ValueError(f"case needs {fp} B of device memory, more than the whole budget {self.budget} B")
        ticket = self._next
        self._next += 1
        self.queue.append(Request(ticket, config, int(steps), tuple(result_fields), fp, time.perf_counter(), prepare=prepare))
        return ticket

    # -- scheduler -----------------------------------------------------------
    def _collect(self):
        tickets = (ctypes.c_int64 * 64)()
        status = (ctypes.c_int32 * 64)()
        n = ctypes.c_int32()
        while True:
            self.lib.lbm_server_poll(self.srv, tickets, status, 64, ctypes.byref(n))
            if n.value == 0:
                return
            for t, st in zip(tickets[: n.value], status[: n.value]):
                req, solver = self.active.pop(int(t))
                try:
                    if st != 0:
                        raise LbmError(f"lbm_step failed ({st}): {solver.case.lib.lbm_last_error(solver.case.h)}")
                    for fid in req.result_fields:
                        req.results[fid] = solver.case.download(fid)
                    req.status = "done"
                except Exception as exc:  # a failing case must not stop the others
                    req.status = "failed"
                    req.error = str(exc)
                req.t_done = time.perf_counter()
                solver.close()
                self.used -= req.footprint
                self.finished.append(req)

    def _admit(self):
        while self.queue and len(self.active) < self.max_active and self.used + self.queue[0].footprint <= self.budget:
            req = self.queue.popleft()
            req.t_admit = time.perf_counter()
            solver = None
            try:
                solver = Solver(config=req.config)
                solver.initialize()
                if req.prepare is not None:
                    req.prepare(solver)
                if self.graphs and solver.case.path() == abi.PATH_FUSED:
                    solver.case.set_option("graph", 1)
                self.used += req.footprint
                self.active[req.ticket] = (req, solver)
                rc = self.lib.lbm_server_add(self.srv, solver.case.h, req.steps, req.ticket)
                if rc != 0:
                    self.active.pop(req.ticket)
                    self.used -= req.footprint
                    raise LbmError(f"lbm_server_add failed ({rc})")
                req.status = "active"
            except Exception as exc:
                req.status = "failed"
                req.error = str(exc)
                req.t_done = time.perf_counter()
                if solver is not None:
                    solver.close()
                self.finished.append(req)
            self.peak_used = max(self.peak_used, self.used)
            self.peak_active = max(self.peak_active, len(self.active))

    def pump(self):
        """One scheduling pass: collect finished cases, admit queued ones."""
        self._collect()
        self._admit()

    def idle(self) -> bool:
        return not self.queue and not self.active

    def run_until_idle(self, poll_s: float = 2e-4):
        while not self.idle():
            self.pump()
            time.sleep(poll_s)

    def close(self):
        if self.srv:
            self.lib.lbm_server_destroy(self.srv)
            self.srv = ctypes.c_void_p()
        for req, solver in self.active.values():
            solver.close()
        self.active.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- statistics ----------------------------------------------------------
    def stats(self, t0: Optional[float] = None) -> dict:
        done = [r for r in self.finished if r.status == "done"]
        if not done:
            return {"cases": 0}
        lat = np.array([r.latency for r in done])
        start = t0 if t0 is not None else min(r.t_submit for r in done)
        wall = max(r.t_done for r in done) - start
        lups = sum(r.cells * r.steps for r in done)
        return {
            "cases": len(done),
            "failed": sum(1 for r in self.finished if r.status == "failed"),
            "wall_s": wall,
            "aggregate_mlups": lups / wall / 1e6,
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "peak_active": self.peak_active,
            "peak_hbm_bytes": self.peak_used,
        }


# --------------------------------------------------------------------------
# the serving stream of the benchmark (sizes, sweeps, step counts)
SIZES = ((32, 32, 32), (64, 32, 32), (96, 48, 48), (128, 64, 64))


def ensemble_case_env(kind: str, size, rng: np.random.Generator, storage: str = "fp32") -> dict:
    """One case of the serving stream: an air case over a moving sinusoidal
    bed (A ~ U[1, 4] cells, wavelength nx/4, mean inflow ~ U[0.03, 0.08]) or a
    water Gaussian mound on a flat pool (amplitude ~ U[0, 8] cells, width
    nx/8), dx = 1 m."""
    nx, ny, nz = size
    if kind == "air":
        from lbm_config import preset_env

        env = preset_env("waves_256x128x128")
        amp = float(rng.uniform(1.0, 4.0))
        env.update({
            "LBM_LX_M": str(nx), "LBM_LY_M": str(ny), "LBM_LZ_M": str(nz),
            "LBM_BED_AMP_M": f"{amp:.6f}", "LBM_BED_WAVELENGTH_M": str(nx / 4.0),
            "LBM_INFLOW_U_LU": f"{float(rng.uniform(0.03, 0.08)):.6f}", "LBM_STORAGE": storage,
        })
        return env
    amp = float(rng.uniform(0.0, 8.0))
    return {
        "LBM_PHYSICS_MODE": "water", "LBM_FREE_SURFACE_MODE": "height_kinematic",
        "LBM_FREE_SURFACE_TRACKING": "vof", "LBM_FREE_SURFACE_INITIAL": "gaussian",
        "LBM_PRESSURE_FORMULATION": "vof_component_balanced", "LBM_COLLISION": "regularized",
        "LBM_USE_LES": "1", "LBM_LES_CS": "0.16", "LBM_TAU": "0.52", "LBM_DX_M": "1",
        "LBM_LX_M": str(nx), "LBM_LY_M": str(ny), "LBM_LZ_M": str(nz),
        "LBM_BED_PROFILE": "flat", "LBM_BED_LEVEL_M": "0", "LBM_FREE_SURFACE_DEPTH_M": str(nz // 2),
        "LBM_GAUSSIAN_AMPLITUDE_M": f"{amp:.6f}", "LBM_GAUSSIAN_SIGMA_M": str(nx / 8.0),
        "LBM_GAUSSIAN_CENTER_X_M": str(nx / 2.0), "LBM_X_BOUNDARY": "solid", "LBM_Y_BOUNDARY": "solid",
        "LBM_VOF_INITIAL_SHAPE": "none", "LBM_G_LB": "1e-5", "LBM_SEED": "7", "LBM_STORAGE": storage,
    }


def ensemble_stream(n_cases: int, seed: int = 7, storage: str = "fp32", steps_range=(500, 2000)):
    """(env, steps) of n_cases requests: half air, half water, sizes uniform
    over SIZES, steps ~ U[steps_range], every case stored in `storage`."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(n_cases):
        kind = "air" if c % 2 == 0 else "water"
        size = SIZES[int(rng.integers(0, len(SIZES)))]
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        env = ensemble_case_env(kind, size, rng, storage)
        out.append((env, steps))
    return out
