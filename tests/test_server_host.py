"""Host side of the ensemble server (no GPU): the serving stream and the
footprint admission bounds."""
import pytest

from lbm_config import parse_env


def test_stream_composition():
    from lbm_server import SIZES, ensemble_stream

    s = ensemble_stream(64, seed=7)
    assert s == ensemble_stream(64, seed=7)
    kinds = [parse_env(env).physics_mode for env, _ in s]
    assert kinds.count("air") == kinds.count("water") == 32
    for env, steps in s:
        c = parse_env(env)
        assert (c.Nx, c.Ny, c.Nz) in SIZES
        assert 500 <= steps <= 2000
        assert not c.unsupported
        if c.physics_mode == "air":
            assert 1.0 <= c.bed_amp_m <= 4.0 and c.bed_wavelength_m == c.Nx / 4
        else:
            assert 0.0 <= c.gaussian_amp_m <= 8.0 and c.gaussian_sigma_m == c.Nx / 8


def test_footprint_and_oversize_rejection():
    from lbm_native import LIB_PATH

    if not LIB_PATH.exists():
        pytest.skip("liblbm.so not built")
    from lbm_server import EnsembleServer, ensemble_case_env, footprint_bytes
    import numpy as np

    rng = np.random.default_rng(0)
    small = parse_env(ensemble_case_env("air", (32, 32, 32), rng))
    big = parse_env(ensemble_case_env("air", (128, 64, 64), rng))
    f16 = parse_env(ensemble_case_env("air", (128, 64, 64), rng, storage="fp16"))
    assert footprint_bytes(small) < footprint_bytes(big)
    assert footprint_bytes(f16) < footprint_bytes(big)
    assert footprint_bytes(big) >= 2 * 19 * 4 * 128 * 64 * 64


def test_oversize_case_is_refused_before_any_device_work():
    from lbm_native import LIB_PATH

    if not LIB_PATH.exists():
        pytest.skip("liblbm.so not built")
    import numpy as np

    from lbm_server import EnsembleServer, ensemble_case_env, footprint_bytes

    rng = np.random.default_rng(0)
    small = ensemble_case_env("air", (32, 32, 32), rng)
    big = ensemble_case_env("air", (128, 64, 64), rng)
    srv = EnsembleServer(hbm_budget_bytes=footprint_bytes(parse_env(small)), max_active=2, chunk_steps=4, n_workers=1)
    with pytest.raises(ValueError, match="more than the whole budget"):
        srv.submit(big, 10)
    assert srv.idle()
    srv.close()
