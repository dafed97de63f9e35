"""Several fused-path cases advanced together by one launch per step
(lbm_step_batch, the ensemble server's merged launch grid): each case ends
bit for bit where lbm_step alone takes it, over calls with a different
step count per case (a case with fewer steps leaves the later launches),
cases of different sizes and bed profiles, fresh and already stepped
cases, FP32 and FP16 storage; handles that cannot share a launch are
refused; and the server with batching gives every case its result alone."""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of
from lbm_native import LbmError, step_batch
from test_gpu_parity import device_case

pytestmark = pytest.mark.gpu

FIELDS = (abi.FIELD_F, abi.FIELD_UX, abi.FIELD_UZ, abi.FIELD_RHO, abi.FIELD_OMEGA)


def waves(nx, ny, nz, amp, **kw):
    return env_of("waves_256x128x128", LBM_LX_M=nx, LBM_LY_M=ny, LBM_LZ_M=nz, LBM_BED_AMP_M=amp,
                  LBM_BED_WAVELENGTH_M=nx / 4, **kw)


GROUPS = {
    "fp32_kbc": [waves(32, 32, 32, 2), waves(64, 32, 32, 3), waves(48, 16, 24, 1.5), waves(37, 5, 11, 1)],
    "fp16_kbc": [waves(32, 32, 32, 2, LBM_STORAGE="fp16"), waves(64, 16, 24, 3, LBM_STORAGE="fp16")],
    "fp32_reg": [env_of("plain_64x32x32"), env_of("plain_64x32x32", LBM_LX_M=32, LBM_LY_M=16, LBM_LZ_M=16)],
}
CALLS = [(16, 16, 7, 16), (16, 3, 16, 16), (1, 16, 16, 9), (16, 16, 16, 16)]


@pytest.mark.parametrize("group", list(GROUPS))
def test_batch_equals_separate_steps(group):
    envs = GROUPS[group]
    alone = [device_case(e)[1] for e in envs]
    batch = [device_case(e)[1] for e in envs]
    alone[0].step(5)  # one case already stepped before it joins
    batch[0].step(5)
    for call in CALLS:
        ns = [call[i % len(call)] for i in range(len(envs))]
        for c, n in zip(alone, ns):
            c.step(n)
        step_batch(batch, ns)
        for i, (a, b) in enumerate(zip(alone, batch)):
            assert b.step_index() == a.step_index()
            for fid in FIELDS:
                x, y = b.download(fid), a.download(fid)
                assert np.array_equal(x, y), (group, call, i, fid)
    for c in alone + batch:
        c.close()


def test_batch_refuses_mixed_cases():
    a = device_case(waves(32, 32, 32, 2))[1]
    b = device_case(waves(32, 32, 32, 2, LBM_STORAGE="fp16"))[1]
    w = device_case(env_of("vof_192x96x160", LBM_LX_M=16, LBM_LY_M=8, LBM_LZ_M=16, LBM_FREE_SURFACE_DEPTH_M=8,
                           LBM_VOF_BLOCK_SIZE_M=4, LBM_VOF_BLOCK_BOTTOM_M=10))[1]
    for cases in ([a, b], [a, w], [a, a]):
        with pytest.raises(LbmError):
            step_batch(cases, [2] * len(cases))
    for c in (a, b, w):
        c.close()


def test_server_with_batching_equals_alone():
    from lbm_server import EnsembleServer
    from test_gpu_server import pocket_case, run_alone, small_stream

    penv, prep = pocket_case()
    stream = [(env, steps, None) for env, steps in small_stream()]
    stream.insert(2, (penv, 9, prep))
    alone = [run_alone(env, steps, p) for env, steps, p in stream]
    srv = EnsembleServer(hbm_budget_bytes=8 << 30, max_active=8, chunk_steps=7, n_workers=2, batch=True)
    tickets = [srv.submit(env, steps, prepare=p) for env, steps, p in stream]
    srv.run_until_idle()
    by_ticket = {r.ticket: r for r in srv.finished}
    assert len(by_ticket) == len(stream)
    for t, ref in zip(tickets, alone):
        r = by_ticket[t]
        assert r.status == "done", r.error
        for fid, arr in ref.items():
            assert np.array_equal(r.results[fid], arr), (t, fid)
    srv.close()
