"""The narrow band of the VOF pipeline (option vof_band, on by default): the
passes skip the rows whose cells are settled far from every change of the
step.  It is an exact shortcut, so a run with the band and a run without it
(every row processed, option vof_band 0) must agree bit for bit in every
field and counter, over runs long enough for the interface to move, cells
to refill, orphans and detached cells to appear and trapped gas to
collapse.  (Parity of the band run with the oracle is covered by the VOF
parity tests, which run with the default.)"""
import numpy as np
import pytest

import lbm_abi as abi
from harness import setup_case
from lbm_config import parse_env, preset_env
from test_gpu_coverage import _droplets, _pocket
from test_oracle_vof import vof_env

pytestmark = pytest.mark.gpu

FIELDS = [abi.FIELD_FS_TYPE, abi.FIELD_FS_FILL, abi.FIELD_FS_MASS, abi.FIELD_F, abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UY,
          abi.FIELD_UZ, abi.FIELD_FS_H, abi.FIELD_OMEGA]


def band_pair(env, prepare=None):
    from lbm_native import Case

    cfg = parse_env(env)
    cases = []
    for band in (1, 0):
        c = Case(cfg.to_struct())
        c.set_option("vof_band", band)
        setup_case(c, cfg)
        if prepare is not None:
            prepare(c, cfg)
        cases.append(c)
    return cfg, cases[0], cases[1]


def same(a, b, what):
    for fid in FIELDS:
        assert np.array_equal(a.download(fid), b.download(fid)), (what, fid)
    assert np.array_equal(a.download(abi.FIELD_VOF_EXTERIOR_GAS), b.download(abi.FIELD_VOF_EXTERIOR_GAS)), what
    assert a.counters() == b.counters(), what


def run(env, chunks, prepare=None, seen_keys=()):
    """Steps both runs one step at a time and compares them after each
    chunk; returns the counts of seen_keys over the steps after the first
    (the first step after the state is set processes every row)."""
    cfg, a, b = band_pair(env, prepare)
    seen = {k: 0 for k in seen_keys}
    for n in chunks:
        for _ in range(n):
            a.step(1)
            b.step(1)
            if a.step_index() > 1:
                cb = b.counters()
                for k in seen_keys:
                    seen[k] += cb[k]
        same(a, b, a.step_index())
    a.close()
    b.close()
    return seen


def test_band_falling_block_with_collapses():
    """Falling block into a pool, 300 steps, trapped gas collapsed every 25
    steps: refills, orphans and collapses happen while the band is on."""
    print("case band: falling block 300 steps", flush=True)
    seen = run(vof_env(LBM_VOF_COLLAPSE_INTERVAL=25), (1, 49, 50, 50, 50, 50, 50),
               seen_keys=("free_surface_refill_count", "free_surface_vof_orphan_count"))
    assert seen["free_surface_refill_count"] > 0
    assert seen["free_surface_vof_orphan_count"] > 0


def test_band_detached_droplets():
    print("case band: detached droplets 12 steps", flush=True)
    seen = run(vof_env(LBM_VOF_INITIAL_SHAPE="none"), (1, 1, 1, 3, 6), prepare=_droplets,
               seen_keys=("free_surface_vof_detached_advect_count", "free_surface_vof_detached_compress_count"))
    assert seen["free_surface_vof_detached_advect_count"] > 0
    assert seen["free_surface_vof_detached_compress_count"] > 0


def test_band_full_size_vof_workload():
    """192 x 96 x 160, the VOF benchmark workload with a gas pocket in the
    pool, collapse every 10 steps, 40 steps."""
    print("case band: vof_192x96x160 40 steps", flush=True)
    env = dict(preset_env("vof_192x96x160"), LBM_VOF_COLLAPSE_INTERVAL="10")
    seen = run(env, (1, 19, 20), prepare=_pocket, seen_keys=("free_surface_vof_collapse_cell_count",))
    assert seen["free_surface_vof_collapse_cell_count"] > 0


@pytest.mark.parametrize("storage", ["fp16", "fp64"])
def test_band_other_storage(storage):
    print("case band: falling block", storage, flush=True)
    run(vof_env(LBM_STORAGE=storage), (1, 39, 40))
