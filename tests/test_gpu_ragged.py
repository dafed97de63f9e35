"""Ragged and minimal grids: x extents that do not fill a block (partial
waves, wrap at x = 0 and nx - 1 inside one wave), odd y and z extents and
the smallest grid the library accepts (2 x 2 x 3).  Both paths and FP16
storage against the oracle, to the same bar as the regular grids."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of
from test_gpu_parity import F32_TOL, device_case, oracle_case, relerr

pytestmark = pytest.mark.gpu

CASES = {
    "waves_37x5x11": env_of("waves_256x128x128", LBM_LX_M=37, LBM_LY_M=5, LBM_LZ_M=11, LBM_BED_AMP_M=1,
                            LBM_BED_WAVELENGTH_M=18.5),
    "plain_100x7x9": env_of("plain_64x32x32", LBM_LX_M=100, LBM_LY_M=7, LBM_LZ_M=9),
    "plain_solidx_67x3x5": env_of("plain_64x32x32", LBM_LX_M=67, LBM_LY_M=3, LBM_LZ_M=5, LBM_X_BOUNDARY="solid",
                                  LBM_COLLISION="kbc"),
    "minimal_2x2x3": env_of("plain_64x32x32", LBM_LX_M=2, LBM_LY_M=2, LBM_LZ_M=3, LBM_BED_LEVEL_M=0.25),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("path,storage", [("staged", "fp32"), ("fused", "fp32"), ("fused", "fp16")])
def test_ragged_grid_parity(name, path, storage):
    env = dict(CASES[name], LBM_KERNEL_PATH=path, LBM_STORAGE=storage)
    print(f"ragged case {name} {path} {storage}", flush=True)
    cfg, o = oracle_case(env)
    _, c = device_case(env)
    assert c.path() == (abi.PATH_STAGED if path == "staged" else abi.PATH_FUSED)
    for chunk in (1, 29):
        o.step(chunk)
        c.step(chunk)
        for fid in (abi.FIELD_F, abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ):
            a, b = c.download(fid), o.download(fid)
            assert np.all(np.isfinite(a)), (name, fid)
            assert relerr(a, b) <= F32_TOL, (name, path, storage, c.step_index(), fid, relerr(a, b))
    for fid in (abi.FIELD_LATTICE_OPEN, abi.FIELD_OBSTACLE):
        assert np.array_equal(c.download(fid), o.download(fid)), (name, fid)
    assert c.counters()["boundary_cut_link_count"] == o.counters()["boundary_cut_link_count"]
    c.close()
    o.close()


def test_grid_below_the_minimum_is_refused():
    from lbm_config import parse_env
    from lbm_native import Case, LbmError

    cfg = parse_env(env_of("plain_64x32x32", LBM_LX_M=2, LBM_LY_M=2, LBM_LZ_M=2))
    with pytest.raises(LbmError):
        Case(cfg.to_struct())
