"""ctypes wrapper of the CPU oracle (oracle/liblbm_oracle.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py as the checker.  The interface mirrors
lbm_native.Case so a test can drive both the same way.
"""
from __future__ import annotations

import ctypes
import subprocess
from pathlib import Path

import numpy as np

import lbm_abi as abi

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "liblbm_oracle.so"

_lib = None


def build_oracle() -> Path:
    res = subprocess.run(["make", "-C", str(ORACLE_DIR)], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stderr)
    return ORACLE_LIB


def load_oracle():
    global _lib
    if _lib is not None:
        return _lib
    if not ORACLE_LIB.exists():
        build_oracle()
    lib = ctypes.CDLL(str(ORACLE_LIB))
    for sfx in ("_f32", "_f64"):
        g = lambda n: getattr(lib, n + sfx)
        g("orc_create").argtypes = [ctypes.POINTER(abi.LbmConfig)]
        g("orc_create").restype = ctypes.c_void_p
        g("orc_destroy").argtypes = [ctypes.c_void_p]
        g("orc_destroy").restype = None
        g("orc_field").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        g("orc_field").restype = ctypes.c_int
        g("orc_geometry").argtypes = [ctypes.c_void_p, ctypes.c_float]
        g("orc_geometry").restype = None
        g("orc_init_populations").argtypes = [ctypes.c_void_p]
        g("orc_init_populations").restype = None
        g("orc_step").argtypes = [ctypes.c_void_p, ctypes.c_int]
        g("orc_step").restype = None
        g("orc_get_counters").argtypes = [ctypes.c_void_p, ctypes.POINTER(abi.LbmCounters)]
        g("orc_get_counters").restype = None
        g("orc_get_step").argtypes = [ctypes.c_void_p]
        g("orc_get_step").restype = ctypes.c_int64
        g("orc_reduce").argtypes = [ctypes.c_void_p, ctypes.c_int]
        g("orc_reduce").restype = ctypes.c_double
        g("orc_stage").argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_float]
        g("orc_stage").restype = ctypes.c_int
        g("orc_field_fnew").argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        g("orc_field_fnew").restype = ctypes.c_int
        g("orc_set_step").argtypes = [ctypes.c_void_p, ctypes.c_int64]
        g("orc_set_step").restype = None
        g("orc_round16").argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
        g("orc_round16").restype = None
        for fn in ("orc_det_sin", "orc_det_cos", "orc_det_log", "orc_det_exp"):
            g(fn).argtypes = [ctypes.c_double]
            g(fn).restype = ctypes.c_double
        g("orc_set_threads").argtypes = [ctypes.c_int]
        g("orc_set_threads").restype = None
        g("orc_get_threads").argtypes = []
        g("orc_get_threads").restype = ctypes.c_int
    _lib = lib
    return lib


_default_threads = None


def set_threads(n: int) -> int:
    """Threads of the oracle's per-cell loops (results do not depend on
    it; n <= 0 restores the count the library started with); returns the
    count now in effect."""
    global _default_threads
    lib = load_oracle()
    if _default_threads is None:
        _default_threads = int(lib.orc_get_threads_f32())
    n = int(n) if n > 0 else _default_threads
    lib.orc_set_threads_f32(n)
    lib.orc_set_threads_f64(n)
    return int(lib.orc_get_threads_f32())


def round16(half: bool, stochastic: bool, x: np.ndarray, r: np.ndarray) -> np.ndarray:
    """The oracle's 16-bit storage rounding of raw float32 values (binary16
    when ``half``, else bfloat16), as exactly representable float32."""
    lib = load_oracle()
    x = np.ascontiguousarray(x, dtype=np.float32)
    r = np.ascontiguousarray(r, dtype=np.uint32)
    out = np.empty_like(x)
    lib.orc_round16_f32(int(half), int(stochastic), x.ctypes.data, r.ctypes.data, out.ctypes.data, x.size)
    return out


class OracleCase:
    """CPU restatement of one case, FP32 (``fp64=False``) or FP64."""

    def __init__(self, cfg: abi.LbmConfig, fp64: bool = False):
        self.lib = load_oracle()
        self.sfx = "_f64" if fp64 else "_f32"
        self.cfg = cfg.copy()
        self.real = np.float64 if fp64 else np.float32
        self.p = self._fn("orc_create")(ctypes.byref(self.cfg))
        if not self.p:
            raise RuntimeError("orc_create rejected the config")

    def _fn(self, name):
        return getattr(self.lib, name + self.sfx)

    def close(self):
        if self.p:
            self._fn("orc_destroy")(self.p)
            self.p = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def dtype_of(self, field):
        return np.int32 if field in abi.INT_FIELDS else self.real

    def upload(self, field, array):
        c = self.cfg
        shape = abi.field_shape(field, c.nx, c.ny, c.nz)
        a = np.ascontiguousarray(array, dtype=self.dtype_of(field)).reshape(shape)
        rc = self._fn("orc_field")(self.p, field, a.ctypes.data, a.nbytes, 0)
        if rc != 0:
            raise RuntimeError(f"oracle upload {field} failed")

    def download(self, field):
        c = self.cfg
        shape = abi.field_shape(field, c.nx, c.ny, c.nz)
        out = np.empty(shape, dtype=self.dtype_of(field))
        rc = self._fn("orc_field")(self.p, field, out.ctypes.data, out.nbytes, 1)
        if rc != 0:
            raise RuntimeError(f"oracle download {field} failed")
        return out

    def download_fnew(self):
        c = self.cfg
        out = np.empty((19, c.nz, c.ny, c.nx), dtype=self.real)
        self._fn("orc_field_fnew")(self.p, out.ctypes.data, out.nbytes)
        return out

    def geometry(self, t=0.0):
        self._fn("orc_geometry")(self.p, ctypes.c_float(t))

    def init_populations(self):
        self._fn("orc_init_populations")(self.p)

    def step(self, n=1):
        self._fn("orc_step")(self.p, int(n))

    def stage(self, stage, t=0.0):
        rc = self._fn("orc_stage")(self.p, int(stage), ctypes.c_float(t))
        if rc != 0:
            raise RuntimeError(f"unknown oracle stage {stage}")

    def set_step(self, step):
        self._fn("orc_set_step")(self.p, int(step))

    def counters(self):
        cnt = abi.LbmCounters()
        self._fn("orc_get_counters")(self.p, ctypes.byref(cnt))
        return cnt.as_dict()

    def reduce(self, what):
        return float(self._fn("orc_reduce")(self.p, int(what)))

    def step_index(self):
        return int(self._fn("orc_get_step")(self.p))
