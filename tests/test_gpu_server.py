"""Ensemble server on the GPU: every case gives the same result alone and in
the mix (bit for bit: the kernels are deterministic), admission honours the
HBM budget, and requests that arrive while others run are admitted as
memory frees up."""
import numpy as np
import pytest

import lbm_abi as abi
from lbm_config import parse_env

pytestmark = pytest.mark.gpu


def small_stream():
    from lbm_server import ensemble_stream

    cases = []
    for storage, seed in (("fp32", 7), ("fp16", 8)):
        for env, _ in ensemble_stream(4, seed=seed, storage=storage, steps_range=(40, 90)):
            cases.append(env)
    rng = np.random.default_rng(3)
    return [(env, int(rng.integers(40, 91))) for env in cases]


def run_alone(env, steps):
    from lbm_solver import Solver

    s = Solver(config=parse_env(env))
    s.initialize()
    s.step(steps)
    out = {fid: s.case.download(fid) for fid in (abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)}
    s.close()
    return out


def test_alone_equals_mix():
    from lbm_server import EnsembleServer

    stream = small_stream()
    alone = [run_alone(env, steps) for env, steps in stream]
    srv = EnsembleServer(hbm_budget_bytes=8 << 30, max_active=8, chunk_steps=7, n_workers=3)
    tickets = [srv.submit(env, steps) for env, steps in stream]
    srv.run_until_idle()
    by_ticket = {r.ticket: r for r in srv.finished}
    assert len(by_ticket) == len(stream)
    assert srv.peak_active > 1
    for t, ref in zip(tickets, alone):
        r = by_ticket[t]
        assert r.status == "done", r.error
        for fid, arr in ref.items():
            assert np.array_equal(r.results[fid], arr), (t, fid)
    st = srv.stats()
    assert st["cases"] == len(stream) and st["failed"] == 0
    assert st["latency_p99_s"] >= st["latency_p50_s"] > 0
    srv.close()


def test_budget_admission_and_late_arrivals():
    from lbm_server import EnsembleServer, ensemble_stream, footprint_bytes

    stream = ensemble_stream(6, seed=11, steps_range=(30, 60))
    fps = [footprint_bytes(parse_env(env)) for env, _ in stream]
    budget = max(fps) + min(fps)
    srv = EnsembleServer(hbm_budget_bytes=budget, max_active=8, chunk_steps=5, n_workers=2)
    for env, steps in stream[:3]:
        srv.submit(env, steps)
    for _ in range(20):
        srv.pump()
    for env, steps in stream[3:]:  # arrive while the first ones run
        srv.submit(env, steps)
    srv.run_until_idle()
    assert srv.peak_used <= budget
    done = [r for r in srv.finished if r.status == "done"]
    assert len(done) == len(stream)
    for r in done:
        assert all(np.all(np.isfinite(a)) for a in r.results.values())
        assert r.t_admit >= r.t_submit and r.t_done >= r.t_admit
    srv.close()
