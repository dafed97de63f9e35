"""Record the rounding of gfx950's v_cvt_sr_f16_f32 / v_cvt_sr_bf16_f32 on
seeded inputs (needs the GPU): writes sr_hw_vectors.npz with x (float32),
r (uint32), f16 and bf16 (uint16 bit patterns).

Inputs: binary16 normal range (|x| in 2^-14 .. 2^16), its subnormal range
and below (2^-27 .. 2^-14), the whole float32 range, overflow neighbours of
65504, zeros, infinities, NaN, with random 32-bit words r plus structured r
(0, all ones, the 13-bit boundaries) on the special values.

usage: python tests/golden/make_sr_vectors.py [out.npz]
"""
import ctypes
import subprocess
import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent


def inputs(seed: int = 2024, per_range: int = 3000):
    rng = np.random.default_rng(seed)
    parts = []
    for lo, hi in ((-14.0, 16.0), (-27.0, -14.0), (-130.0, 128.0)):
        e = rng.uniform(lo, hi, per_range)
        parts.append((2.0**e) * rng.choice([-1.0, 1.0], per_range))
    near = np.array([65504.0, 65505.0, 65519.0, 65520.0, 65535.0, 65536.0, 2.0**-14, 2.0**-24, 2.0**-25,
                     1.5 * 2.0**-24, 3.0 * 2.0**-26, 1.0, -1.0, 0.1, -0.1], np.float64)
    x = np.concatenate(parts + [near]).astype(np.float32)
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan], np.float32)
    rs = np.array([0, 1, 0x7FFFF, 0x80000, 0xFFF80000, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFF], np.uint32)
    r = rng.integers(0, 2**32, x.size, dtype=np.uint64).astype(np.uint32)
    x = np.concatenate([x, np.repeat(near.astype(np.float32), rs.size), np.repeat(special, rs.size)])
    r = np.concatenate([r, np.tile(rs, near.size), np.tile(rs, special.size)])
    return x, r


def main(out: Path):
    x, r = inputs()
    with tempfile.TemporaryDirectory() as td:
        so = Path(td) / "libsrprobe.so"
        cmd = ["hipcc", "--offload-arch=gfx950", "-O2", "-shared", "-fPIC", str(HERE / "sr_probe.hip"), "-o", str(so)]
        subprocess.run(cmd, check=True)
        lib = ctypes.CDLL(str(so))
        o16 = np.zeros(x.size, np.uint16)
        ob = np.zeros(x.size, np.uint16)
        p = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        rc = lib.sr_probe(p(x), p(r), p(o16), p(ob), ctypes.c_int(x.size))
        if rc != 0:
            raise RuntimeError(f"sr_probe failed ({rc})")
    np.savez_compressed(out, x=x, r=r, f16=o16, bf16=ob)
    print(f"wrote {out}: {x.size} vectors")


if __name__ == "__main__":
    main(Path(sys.argv[1]) if len(sys.argv) > 1 else HERE / "sr_hw_vectors.npz")
