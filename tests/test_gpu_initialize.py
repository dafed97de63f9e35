"""lbm_initialize (the whole init sequence of CODEX:3777-3792 inside the
library, for callers without Python) against Solver.initialize() (geometry,
the Python host fields, their upload and lbm_init_populations): every field
the handle holds is bit-identical after the init and after a few steps, on
air and water cases, both kernel paths and the storage formats."""
import numpy as np
import pytest

import lbm_abi as abi
from lbm_config import parse_env
from lbm_native import Case, LbmError
from lbm_solver import Solver
from test_init_native import BOTTOM, CASES

pytestmark = pytest.mark.gpu

RUNS = [
    ("uniform_noise", {}),
    ("uniform_noise", {"LBM_STORAGE": "fp16", "LBM_KERNEL_PATH": "fused"}),
    ("reference_kick", {"LBM_KERNEL_PATH": "staged"}),
    ("reference_kick_fp64", {}),
    ("loglaw_inflow", {"LBM_STORAGE": "bf16"}),
    ("moving_bed", {}),
    ("water_solitary_divfree", {}),
    ("water_gauss_linear_wave_solid", {}),
    ("water_vof_block", {}),
    ("water_piecewise_bed", {}),
]


def fields_of(case):
    out = {}
    for fid in range(abi.FIELD_COUNT):
        try:
            out[fid] = case.download(fid)
        except (LbmError, ValueError):
            pass
    return out


@pytest.mark.parametrize("name,extra", RUNS, ids=[f"{n}-{'-'.join(e.values()) or 'default'}" for n, e in RUNS])
def test_initialize_matches_solver(name, extra, tmp_path):
    env = dict(CASES[name], **extra)
    if env.get("LBM_BED_PROFILE") == "piecewise_linear":
        (tmp_path / "bottom.txt").write_text(BOTTOM)
        env["LBM_BOTTOM_FILE"] = str(tmp_path / "bottom.txt")
    print("case", name, extra, flush=True)
    cfg = parse_env(env)
    solver = Solver(config=cfg)
    solver.initialize()
    c = Case(cfg.to_struct())
    if cfg.bed_profile == "piecewise_linear":
        c.upload(abi.FIELD_PIECEWISE_BED_Z, cfg.piecewise_bed_z_np)
    c.initialize(cfg.init_params())
    for steps in (0, 3):
        if steps:
            solver.step(steps)
            c.step(steps)
        want, got = fields_of(solver.case), fields_of(c)
        assert set(want) == set(got) and abi.FIELD_F in got
        for fid, a in want.items():
            assert np.array_equal(a, got[fid], equal_nan=True), (name, steps, fid, int(np.sum(a != got[fid])))
        assert solver.counters() == c.counters()


def test_initialize_rejects_a_bad_version():
    cfg = parse_env(CASES["uniform_noise"])
    c = Case(cfg.to_struct())
    p = cfg.init_params()
    p.version = 0
    with pytest.raises(LbmError, match="version"):
        c.initialize(p)
