"""Ensemble server on the GPU: every case gives the same result alone and in
the mix (bit for bit: the kernels are deterministic, a trapped-gas collapse
included), admission honours the HBM budget, requests that arrive while
others run are admitted as memory frees up, and one handle cannot be handed
to the server twice."""
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


def pocket_case():
    """A water case with a sealed gas pocket and the trapped-gas collapse
    on every step: its FP64 candidate / capacity sums scale the state."""
    from test_oracle_vof import _pocket_state, vof_env

    env = vof_env(LBM_VOF_INITIAL_SHAPE="none", LBM_VOF_COLLAPSE_INTERVAL=1, LBM_VOF_COLLAPSE_MAX_VOLUME_M3=64)
    cfg = parse_env(env)

    def prepare(solver):
        fill, ctype = _pocket_state(cfg, ((8, 11), (6, 9), (10, 13)))
        rho = solver.case.download(abi.FIELD_RHO)
        mass = np.where(ctype == abi.FS_GAS, 0.0, fill * rho).astype(np.float32)
        solver.case.upload(abi.FIELD_FS_FILL, fill)
        solver.case.upload(abi.FIELD_FS_TYPE, ctype)
        solver.case.upload(abi.FIELD_FS_MASS, mass)

    return env, prepare


def run_alone(env, steps, prepare=None):
    from lbm_solver import Solver

    s = Solver(config=parse_env(env))
    s.initialize()
    if prepare is not None:
        prepare(s)
    s.step(steps)
    out = {fid: s.case.download(fid) for fid in (abi.FIELD_RHO, abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)}
    s.close()
    return out


def test_alone_equals_mix():
    from lbm_server import EnsembleServer

    penv, prep = pocket_case()
    stream = [(env, steps, None) for env, steps in small_stream()]
    stream.insert(3, (penv, 9, prep))
    alone = [run_alone(env, steps, p) for env, steps, p in stream]
    srv = EnsembleServer(hbm_budget_bytes=8 << 30, max_active=8, chunk_steps=7, n_workers=3)
    tickets = [srv.submit(env, steps, prepare=p) for env, steps, p in stream]
    srv.run_until_idle()
    by_ticket = {r.ticket: r for r in srv.finished}
    assert len(by_ticket) == len(stream)
    assert srv.peak_active > 1
    for t, ref in zip(tickets, alone):
        r = by_ticket[t]
        assert r.status == "done", r.error
        for fid, arr in ref.items():
            assert np.array_equal(r.results[fid], arr), (t, fid)
    # the pocket case really collapsed (its candidate sums decided a removal)
    from lbm_solver import Solver

    s = Solver(config=parse_env(penv))
    s.initialize()
    prep(s)
    s.step(1)
    assert s.case.counters()["free_surface_vof_collapse_cell_count"] > 0
    s.close()
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


def test_server_refuses_a_handle_it_holds():
    from lbm_native import load_library
    from lbm_solver import Solver
    import ctypes

    lib = load_library()
    srv = ctypes.c_void_p()
    assert lib.lbm_server_create(4, 1, ctypes.byref(srv)) == 0
    env, _ = small_stream()[0]
    s = Solver(config=parse_env(env))
    s.initialize()
    assert lib.lbm_server_add(srv, s.case.h, 200, 1) == 0
    assert lib.lbm_server_add(srv, s.case.h, 200, 2) == abi.LBM_E_STATE
    n = ctypes.c_int32()
    tickets = (ctypes.c_int64 * 4)()
    status = (ctypes.c_int32 * 4)()
    got = []
    import time
    t0 = time.time()
    while not got and time.time() - t0 < 60:
        lib.lbm_server_poll(srv, tickets, status, 4, ctypes.byref(n))
        got += [(tickets[i], status[i]) for i in range(n.value)]
        time.sleep(1e-3)
    assert got == [(1, 0)]
    assert lib.lbm_server_add(srv, s.case.h, 3, 3) == 0  # free again once finished
    assert lib.lbm_server_destroy(srv) == 0
    s.close()
