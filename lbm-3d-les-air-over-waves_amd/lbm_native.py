"""ctypes binding of liblbm.so (the C ABI declared in include/lbm.h).

``Case`` wraps one lbm_handle_t.  There is no fallback: when the library is
missing or fails to load, constructing a Case raises.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Optional

import numpy as np

import lbm_abi as abi

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "liblbm.so"

_lib = None

EXPORTS = [
    "lbm_create",
    "lbm_destroy",
    "lbm_geometry",
    "lbm_init_populations",
    "lbm_initialize",
    "lbm_initial_fields",
    "lbm_step",
    "lbm_step_batch",
    "lbm_upload_field",
    "lbm_download_field",
    "lbm_get_counters",
    "lbm_reduce",
    "lbm_reduce_noalloc",
    "lbm_get_step",
    "lbm_sync",
    "lbm_last_error",
    "lbm_last_step_timing",
    "lbm_device_bytes",
    "lbm_set_option",
    "lbm_get_option",
    "lbm_get_path",
    "lbm_query",
    "lbm_server_create",
    "lbm_server_batch",
    "lbm_server_add",
    "lbm_server_poll",
    "lbm_server_active",
    "lbm_server_destroy",
]


def load_library(path: Optional[os.PathLike] = None) -> ctypes.CDLL:
    """Load liblbm.so once and declare the prototypes."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise RuntimeError(f"{p} is missing: build it with __graft_entry__.build() (hipcc, gfx950)")
    lib = ctypes.CDLL(str(p))
    H = ctypes.c_void_p
    lib.lbm_create.argtypes = [ctypes.POINTER(abi.LbmConfig), ctypes.POINTER(H)]
    lib.lbm_create.restype = ctypes.c_int
    lib.lbm_destroy.argtypes = [H]
    lib.lbm_destroy.restype = None
    lib.lbm_geometry.argtypes = [H, ctypes.c_float]
    lib.lbm_init_populations.argtypes = [H]
    lib.lbm_initialize.argtypes = [H, ctypes.POINTER(abi.LbmInitParams)]
    P = ctypes.c_void_p
    lib.lbm_initial_fields.argtypes = [ctypes.POINTER(abi.LbmConfig), ctypes.POINTER(abi.LbmInitParams), P, P, P, P, P, P,
                                       P, P, P, P, P]
    lib.lbm_step.argtypes = [H, ctypes.c_int32]
    lib.lbm_step_batch.argtypes = [ctypes.POINTER(H), ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    lib.lbm_upload_field.argtypes = [H, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]
    lib.lbm_download_field.argtypes = [H, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t]
    lib.lbm_get_counters.argtypes = [H, ctypes.POINTER(abi.LbmCounters)]
    lib.lbm_reduce.argtypes = [H, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    lib.lbm_reduce_noalloc.argtypes = [H, ctypes.c_int32, ctypes.POINTER(ctypes.c_double)]
    lib.lbm_get_step.argtypes = [H, ctypes.POINTER(ctypes.c_int64)]
    lib.lbm_sync.argtypes = [H]
    lib.lbm_last_error.argtypes = [H]
    lib.lbm_last_error.restype = ctypes.c_char_p
    lib.lbm_last_step_timing.argtypes = [H, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    lib.lbm_device_bytes.argtypes = [ctypes.POINTER(abi.LbmConfig), ctypes.POINTER(ctypes.c_size_t)]
    lib.lbm_set_option.argtypes = [H, ctypes.c_char_p, ctypes.c_int64]
    lib.lbm_get_option.argtypes = [H, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64)]
    lib.lbm_get_path.argtypes = [H, ctypes.POINTER(ctypes.c_int32)]
    lib.lbm_query.argtypes = [H, ctypes.POINTER(ctypes.c_int32)]
    S = ctypes.c_void_p
    lib.lbm_server_create.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(S)]
    lib.lbm_server_batch.argtypes = [S, ctypes.c_int32]
    lib.lbm_server_add.argtypes = [S, H, ctypes.c_int32, ctypes.c_int64]
    lib.lbm_server_poll.argtypes = [S, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32), ctypes.c_int32,
                                    ctypes.POINTER(ctypes.c_int32)]
    lib.lbm_server_active.argtypes = [S, ctypes.POINTER(ctypes.c_int32)]
    lib.lbm_server_destroy.argtypes = [S]
    for name in EXPORTS:
        if name not in ("lbm_destroy", "lbm_last_error"):
            getattr(lib, name).restype = ctypes.c_int
    if path is None:
        _lib = lib
    return lib


class LbmError(RuntimeError):
    pass


def step_batch(cases, nsteps):
    """lbm_step_batch: advance case i by nsteps[i] steps, every step of all
    of them one launch (fused path; one storage type and collision model)."""
    lib = load_library()
    n = len(cases)
    hs = (ctypes.c_void_p * n)(*[c.h for c in cases])
    ns = (ctypes.c_int32 * n)(*[int(x) for x in nsteps])
    rc = lib.lbm_step_batch(hs, ns, n)
    if rc != 0:
        raise LbmError(f"lbm_step_batch failed ({rc}): {lib.lbm_last_error(cases[0].h).decode()}")


def device_bytes(cfg: abi.LbmConfig) -> int:
    lib = load_library()
    out = ctypes.c_size_t(0)
    rc = lib.lbm_device_bytes(ctypes.byref(cfg), ctypes.byref(out))
    if rc != 0:
        raise LbmError(f"lbm_device_bytes: config rejected ({rc})")
    return int(out.value)


def initial_fields_native(cfg: abi.LbmConfig, params: abi.LbmInitParams, obstacle0: np.ndarray,
                          bed: Optional[np.ndarray] = None) -> dict:
    """lbm_initial_fields: the host initial fields, computed by the library
    without a device, keyed by field id like lbm_init.initial_fields."""
    lib = load_library()
    nx, ny, nz = cfg.nx, cfg.ny, cfg.nz
    real = np.float64 if cfg.storage == abi.STORE_FP64 else np.float32
    obst = np.ascontiguousarray(obstacle0, dtype=np.int32).reshape(nz, ny, nx)
    bed_a = None if bed is None else np.ascontiguousarray(bed, dtype=real).reshape(ny, nx)
    out = {f: np.empty((nz, ny, nx), real) for f in (abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)}
    out[abi.FIELD_U_INLET_PROFILE] = np.empty(nz, real)
    out[abi.FIELD_INFLOW_UX] = np.empty((nz, ny), real)
    water = cfg.fs_enabled != 0
    if water:
        out[abi.FIELD_FS_FILL] = np.empty((nz, ny, nx), real)
        out[abi.FIELD_FS_TYPE] = np.empty((nz, ny, nx), np.int32)
        out[abi.FIELD_FS_H] = np.empty((ny, nx), real)

    def ptr(f):
        return out[f].ctypes.data if f in out else None

    rc = lib.lbm_initial_fields(ctypes.byref(cfg), ctypes.byref(params), obst.ctypes.data,
                                None if bed_a is None else bed_a.ctypes.data, ptr(abi.FIELD_RHO), ptr(abi.FIELD_UX),
                                ptr(abi.FIELD_UY), ptr(abi.FIELD_UZ), ptr(abi.FIELD_U_INLET_PROFILE),
                                ptr(abi.FIELD_INFLOW_UX), ptr(abi.FIELD_FS_FILL), ptr(abi.FIELD_FS_TYPE),
                                ptr(abi.FIELD_FS_H))
    if rc != 0:
        raise LbmError(f"lbm_initial_fields failed ({rc})")
    if water:
        out[abi.FIELD_FS_TYPE_PREV] = out[abi.FIELD_FS_TYPE].copy()
    return out


class Case:
    """One simulation case on the GPU (one lbm_handle_t)."""

    def __init__(self, cfg: abi.LbmConfig):
        self.lib = load_library()
        self.cfg = cfg.copy()
        self.h = ctypes.c_void_p()
        rc = self.lib.lbm_create(ctypes.byref(self.cfg), ctypes.byref(self.h))
        if rc != 0:
            msg = self.lib.lbm_last_error(None)
            raise LbmError(f"lbm_create failed ({rc}): {msg.decode() if msg else ''}")
        self.real = np.float64 if cfg.storage == abi.STORE_FP64 else np.float32

    # -- plumbing ------------------------------------------------------------
    def _check(self, rc: int, what: str):
        if rc != 0:
            msg = self.lib.lbm_last_error(self.h)
            raise LbmError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self.lib.lbm_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- API -----------------------------------------------------------------
    def geometry(self, t: float = 0.0):
        self._check(self.lib.lbm_geometry(self.h, ctypes.c_float(t)), "lbm_geometry")

    def init_populations(self):
        self._check(self.lib.lbm_init_populations(self.h), "lbm_init_populations")

    def initialize(self, params: abi.LbmInitParams):
        """lbm_initialize: the whole init sequence inside the library."""
        self._check(self.lib.lbm_initialize(self.h, ctypes.byref(params)), "lbm_initialize")

    def step(self, n: int = 1):
        self._check(self.lib.lbm_step(self.h, int(n)), "lbm_step")

    def sync(self):
        self._check(self.lib.lbm_sync(self.h), "lbm_sync")

    def dtype_of(self, field: int):
        return np.int32 if field in abi.INT_FIELDS else self.real

    def upload(self, field: int, array: np.ndarray):
        c = self.cfg
        shape = abi.field_shape(field, c.nx, c.ny, c.nz)
        a = np.ascontiguousarray(array, dtype=self.dtype_of(field)).reshape(shape)
        self._check(self.lib.lbm_upload_field(self.h, field, a.ctypes.data, a.nbytes), f"upload field {field}")

    def download(self, field: int) -> np.ndarray:
        c = self.cfg
        shape = abi.field_shape(field, c.nx, c.ny, c.nz)
        out = np.empty(shape, dtype=self.dtype_of(field))
        self._check(self.lib.lbm_download_field(self.h, field, out.ctypes.data, out.nbytes), f"download field {field}")
        return out

    def counters(self) -> dict:
        cnt = abi.LbmCounters()
        self._check(self.lib.lbm_get_counters(self.h, ctypes.byref(cnt)), "lbm_get_counters")
        d = cnt.as_dict()
        d["flood_fill_unconverged"] = cnt.pad0
        return d

    def reduce(self, what: int) -> float:
        out = ctypes.c_double(0.0)
        self._check(self.lib.lbm_reduce(self.h, what, ctypes.byref(out)), "lbm_reduce")
        return float(out.value)

    def reduce_noalloc(self, what: int) -> float:
        """lbm_reduce with the partial sums in the library's static device
        scratch (allocates nothing)."""
        out = ctypes.c_double(0.0)
        self._check(self.lib.lbm_reduce_noalloc(self.h, what, ctypes.byref(out)), "lbm_reduce_noalloc")
        return float(out.value)

    def step_index(self) -> int:
        out = ctypes.c_int64(0)
        self._check(self.lib.lbm_get_step(self.h, ctypes.byref(out)), "lbm_get_step")
        return int(out.value)

    def set_option(self, key: str, value: int):
        self._check(self.lib.lbm_set_option(self.h, key.encode(), int(value)), f"lbm_set_option({key})")

    def get_option(self, key: str) -> int:
        out = ctypes.c_int64(0)
        self._check(self.lib.lbm_get_option(self.h, key.encode(), ctypes.byref(out)), f"lbm_get_option({key})")
        return int(out.value)

    def path(self) -> int:
        out = ctypes.c_int32(0)
        self._check(self.lib.lbm_get_path(self.h, ctypes.byref(out)), "lbm_get_path")
        return int(out.value)

    def last_timing(self):
        a, b = ctypes.c_double(0.0), ctypes.c_double(0.0)
        self._check(self.lib.lbm_last_step_timing(self.h, ctypes.byref(a), ctypes.byref(b)), "lbm_last_step_timing")
        return float(a.value), float(b.value)
