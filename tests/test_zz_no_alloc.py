"""lbm_reduce_noalloc (the reduction entry point that allocates nothing)
against lbm_reduce, through the C ABI: every reduction, on a post-collision
state (fused path, streamed on the fly), on a post-stream state (staged
path) and with 16-bit storage; two calls in a row on one GPU, the second on
a smaller grid, so that the static scratch's leftovers of the first call
cannot leak into the second."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of, setup_case
from lbm_config import parse_env

pytestmark = pytest.mark.gpu

WHATS = [abi.REDUCE_MASS, abi.REDUCE_MAX_SPEED, abi.REDUCE_NONFINITE]


def make(env, steps):
    from lbm_native import Case

    cfg = parse_env(env)
    c = Case(cfg.to_struct())
    setup_case(c, cfg)
    c.step(steps)
    return c


@pytest.mark.parametrize("storage,path", [("fp32", "fused"), ("fp32", "staged"), ("fp16", "fused")])
def test_noalloc_equals_reduce_two_calls(storage, path):
    big = env_of("waves_256x128x128", LBM_LX_M=128, LBM_LY_M=64, LBM_LZ_M=48, LBM_BED_AMP_M=2, LBM_BED_WAVELENGTH_M=32,
                 LBM_STORAGE=storage, LBM_KERNEL_PATH=path)
    small = env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2, LBM_BED_WAVELENGTH_M=32,
                   LBM_STORAGE=storage, LBM_KERNEL_PATH=path)
    print("case", storage, path, flush=True)
    for env in (big, small):  # second call: a smaller input
        c = make(env, 7)
        for what in WHATS:
            a = c.reduce(what)
            b = c.reduce_noalloc(what)
            assert a == b, (storage, path, what, a, b)  # bit for bit
            assert np.isfinite(a)
        c.close()


def test_noalloc_vof_sums():
    from test_oracle_vof import vof_env

    c = make(vof_env(), 5)
    for what in (abi.REDUCE_MASS, abi.REDUCE_VOF_MASS, abi.REDUCE_VOF_VOLUME, abi.REDUCE_MAX_SPEED):
        assert c.reduce(what) == c.reduce_noalloc(what), what
    c.close()
