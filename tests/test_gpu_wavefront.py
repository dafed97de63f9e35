"""The wavefront schedule of the fused step (both passes in one launch, the
collision of a tile waiting on the moments tiles around it) and the
tile-resident form (persistent blocks that run a tile's moments and
collision with the populations kept in registers) against the
three-launch schedule: the same arithmetic in another order of tiles, so
the states agree bit for bit, on every collision model and on ragged grids.
FP32 storage only: the 16-bit storage forms keep the three launches (the
options are accepted and have no effect there, which the fp16 / bf16 cases
check)."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of, setup_case
from lbm_config import parse_env

pytestmark = pytest.mark.gpu

SMALL = dict(LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2, LBM_BED_WAVELENGTH_M=32)
CASES = {
    "waves_kbc": env_of("waves_256x128x128", **SMALL),
    "waves_reg": env_of("waves_256x128x128", LBM_COLLISION="regularized", **SMALL),
    "plain_sponge": env_of("plain_64x32x32", LBM_COLLISION="kbc", LBM_ENABLE_TOP_SPONGE=1, LBM_TOP_BOUNDARY="keep",
                           LBM_LX_M=32, LBM_LY_M=16, LBM_LZ_M=16),
    "ragged": env_of("waves_256x128x128", LBM_LX_M=70, LBM_LY_M=13, LBM_LZ_M=19, LBM_BED_AMP_M=2, LBM_BED_WAVELENGTH_M=35),
    "waves_fp16": env_of("waves_256x128x128", LBM_STORAGE="fp16", **SMALL),
    "waves_bf16": env_of("waves_256x128x128", LBM_STORAGE="bf16", **SMALL),
}


def run(env, steps, **opts):
    from lbm_native import Case

    cfg = parse_env(dict(env, LBM_KERNEL_PATH="fused"))
    c = Case(cfg.to_struct())
    for k, v in opts.items():
        c.set_option(k, v)
    setup_case(c, cfg)
    c.step(steps)
    out = {fid: c.download(fid) for fid in (abi.FIELD_F, abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ, abi.FIELD_RHO, abi.FIELD_OMEGA)}
    c.close()
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_wavefront_equals_three_launches(name):
    print("case", name, flush=True)
    env = CASES[name]
    ref = run(env, 30, wavefront=0)
    for lag in (1, 8, 64):
        got = run(env, 30, wavefront=1, wave_lag=lag, resident=0)
        for fid, a in ref.items():
            assert np.array_equal(got[fid], a), (name, lag, fid, int(np.sum(got[fid] != a)))


@pytest.mark.parametrize("block_x", [64, 128, 256])
def test_wavefront_tile_shapes(block_x):
    """Tiles of 64 x 4, 128 x 2 and 256 x 1 cells: several tiles per row and
    per plane, so the flags of x- and y-neighbouring tiles (wrapped at the
    edges) are what the collision waits on."""
    env = CASES["ragged"]
    ref = run(env, 20, wavefront=0, block_x=block_x)
    got = run(env, 20, wavefront=1, wave_lag=2, block_x=block_x, resident=0)
    for fid, a in ref.items():
        assert np.array_equal(got[fid], a), (block_x, fid, int(np.sum(got[fid] != a)))


def test_wavefront_counters_reset_between_calls():
    """Several lbm_step calls of different lengths in a row: each launch must
    find the counters it needs at zero (the last tile resets them)."""
    env = CASES["waves_kbc"]
    ref = run(env, 1 + 2 + 7, wavefront=0)
    from lbm_native import Case

    cfg = parse_env(dict(env, LBM_KERNEL_PATH="fused"))
    c = Case(cfg.to_struct())
    setup_case(c, cfg)
    for n in (1, 2, 7):
        c.step(n)
        c.sync()
    f = c.download(abi.FIELD_F)
    c.close()
    assert np.array_equal(f, ref[abi.FIELD_F])


@pytest.mark.parametrize("block_x,lx", [(64, 64), (128, 128), (256, 256)])
def test_wavefront_xcd_strips(block_x, lx):
    """Rows of tiles a multiple of 8 (32 rows: 8, 16 and 32 rows of tiles),
    so the tiles of a plane are dealt to the XCDs in strips: the strip order
    and the plain order give the three-launch state bit for bit, at a lag
    below and above the tiles of a plane."""
    env = env_of("waves_256x128x128", LBM_LX_M=lx, LBM_LY_M=32, LBM_LZ_M=20, LBM_BED_AMP_M=2, LBM_BED_WAVELENGTH_M=32)
    print("case strips", block_x, lx, flush=True)
    ref = run(env, 20, wavefront=0, block_x=block_x)
    for strips in (0, 1):
        for lag in (1, 0, 5):
            got = run(env, 20, wavefront=1, wave_lag=lag, block_x=block_x, xcd_strips=strips, resident=0)
            for fid, a in ref.items():
                assert np.array_equal(got[fid], a), (block_x, strips, lag, fid, int(np.sum(got[fid] != a)))


@pytest.mark.parametrize("name", list(CASES))
def test_resident_equals_three_launches(name):
    env = CASES[name]
    ref = run(env, 30, wavefront=0)
    got = run(env, 30, wavefront=1, resident=1)
    for fid, a in ref.items():
        assert np.array_equal(got[fid], a), (name, fid, int(np.sum(got[fid] != a)))


@pytest.mark.parametrize("block_x", [64, 128, 256])
def test_resident_tile_shapes_and_calls(block_x):
    """Tiles of 64 x 4, 128 x 2 and 256 x 1 cells, and several calls in a
    row (the ticket counter carries on from launch to launch): the state
    of three launches bit for bit."""
    env = env_of("waves_256x128x128", LBM_LX_M=256, LBM_LY_M=64, LBM_LZ_M=24, LBM_BED_AMP_M=2, LBM_BED_WAVELENGTH_M=64)
    ref = run(env, 1 + 2 + 7, wavefront=0, block_x=block_x)
    from lbm_native import Case

    cfg = parse_env(dict(env, LBM_KERNEL_PATH="fused"))
    c = Case(cfg.to_struct())
    c.set_option("block_x", block_x)
    c.set_option("resident", 1)
    setup_case(c, cfg)
    for n in (1, 2, 7):
        c.step(n)
    for fid, a in ref.items():
        got = c.download(fid)
        assert np.array_equal(got, a), (block_x, fid, int(np.sum(got != a)))
    c.close()


@pytest.mark.parametrize("name", ["waves_kbc", "waves_fp16"])
def test_schedule_chosen_by_timing(name):
    """Option resident 2 (the default): the first call of at least 18 steps
    times both schedules and keeps the faster; the state is the one of the
    wavefront schedule (resident 0) bit for bit, and a later call keeps
    the choice."""
    from lbm_native import Case

    print("case", name, flush=True)
    env = CASES[name]
    ref = run(env, 40, resident=0)
    cfg = parse_env(dict(env, LBM_KERNEL_PATH="fused"))
    c = Case(cfg.to_struct())
    assert c.get_option("resident") == 2
    setup_case(c, cfg)
    c.step(30)
    chosen = c.get_option("resident")
    assert chosen in (0, 1)
    c.step(10)
    assert c.get_option("resident") == chosen
    for fid, a in ref.items():
        got = c.download(fid)
        assert np.array_equal(got, a), (name, chosen, fid, int(np.sum(got != a)))
    c.close()


def test_handles_stepping_at_once_are_chained():
    """Four handles on the wavefront schedule stepped at the same time from
    four threads (their waiting launches chained by the library): each
    state equals the same case stepped alone, bit for bit, and no wait
    runs out."""
    import threading

    from lbm_native import Case

    env = CASES["waves_kbc"]
    ref = run(env, 48, resident=0)
    cfg = parse_env(dict(env, LBM_KERNEL_PATH="fused"))
    cases = []
    for _ in range(4):
        c = Case(cfg.to_struct())
        c.set_option("resident", 0)
        setup_case(c, cfg)
        cases.append(c)
    errors = []

    def work(c):
        try:
            for _ in range(12):
                c.step(4)
            c.sync()
        except Exception as exc:  # reported below
            errors.append(str(exc))

    threads = [threading.Thread(target=work, args=(c,)) for c in cases]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for c in cases:
        for fid, a in ref.items():
            assert np.array_equal(c.download(fid), a), fid
        c.close()
