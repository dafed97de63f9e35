"""HIP-graph launch of the fused step (option "graph"): every lbm_step call
of several steps is captured into a graph and launched as one, the kept
executable graph updated with each call's new arguments (step number, bed
time levels, tile-flag epoch).  The state after every call equals the one
of direct launches bit for bit, over calls of different lengths (so the
graph is updated, re-instantiated when the launch sequence changes, and
bypassed for single steps), with downloads and reductions between the
calls; and the ensemble server with graphs gives every case the result it
has alone."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of
from test_gpu_parity import device_case

pytestmark = pytest.mark.gpu

CASES = {
    "waves_kbc": env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2,
                        LBM_BED_WAVELENGTH_M=32),
    "plain_reg": env_of("plain_64x32x32"),
    "waves_fp16": env_of("waves_256x128x128", LBM_LX_M=64, LBM_LY_M=16, LBM_LZ_M=24, LBM_BED_AMP_M=2,
                         LBM_BED_WAVELENGTH_M=32, LBM_STORAGE="fp16"),
}
FIELDS = (abi.FIELD_F, abi.FIELD_UX, abi.FIELD_UZ, abi.FIELD_RHO, abi.FIELD_OMEGA)


@pytest.mark.parametrize("name", list(CASES))
def test_graph_calls_equal_direct_launches(name):
    _, d = device_case(CASES[name])
    _, g = device_case(CASES[name])
    assert g.path() == abi.PATH_FUSED
    g.set_option("graph", 1)
    for n in (16, 16, 16, 7, 1, 16, 9, 16, 16):
        d.step(n)
        g.step(n)
        for fid in FIELDS:
            a, b = g.download(fid), d.download(fid)
            assert np.array_equal(a, b), (name, n, g.step_index(), fid)
        assert g.reduce(abi.REDUCE_MASS) == d.reduce(abi.REDUCE_MASS)
    d.close()
    g.close()


def test_server_with_graphs_equals_alone():
    from lbm_config import parse_env
    from lbm_server import EnsembleServer
    from test_gpu_server import run_alone, small_stream

    stream = small_stream()
    alone = [run_alone(env, steps) for env, steps in stream]
    srv = EnsembleServer(hbm_budget_bytes=8 << 30, max_active=8, chunk_steps=16, n_workers=3, graphs=True)
    tickets = [srv.submit(env, steps) for env, steps in stream]
    srv.run_until_idle()
    by_ticket = {r.ticket: r for r in srv.finished}
    assert any(parse_env(env).physics_mode == "air" for env, _ in stream)
    for t, ref in zip(tickets, alone):
        r = by_ticket[t]
        assert r.status == "done", r.error
        for fid, arr in ref.items():
            assert np.array_equal(r.results[fid], arr), (t, fid)
    srv.close()
