"""Pins of the oracle's VOF kernels on hand-built states, each against a
closed form or an exact invariant (FP64 oracle, one kernel at a time):
  * compute_vof_mass_exchange (CODEX:1305-1350): antisymmetry of the
    interface-interface exchange (the sum over a closed box is zero), and on
    a population field linear in x the net exchange of an interface cell
    among liquid neighbours is minus the x flux of its slope;
  * remove_orphan_vof_interface_cells (CODEX:1659-1776): an isolated wet
    cell becomes gas and its mass lands on the resolved wet cells 2..r cells
    away along the 18 rays, weighted by room / d^2 (fill / d^2 when nothing
    has room);
  * advect_detached_vof_cells (CODEX:1779-1914): mass m u dt leaves through
    each face with a gas neighbour downstream, scaled to a total of 0.95 m
    when the rates sum above 0.95;
  * compress_detached_vof_residuals (CODEX:1917-1993): a residual cell
    sends all its mass to the gas neighbour of the largest rate;
  * refill_new_free_surface_cells (CODEX:2163-2316): density and velocity
    of a refilled cell are the fill- and weight-averaged values of its wet
    donors (equilibrium mode), plus the averaged non-equilibrium (noneq),
    or the donors' own populations rescaled to that density (directional);
  * repair_vof_coalescence_links (CODEX:2319-2373): f[q] <- f_src[q] on the
    links whose source was gas and is wet now, moments recomputed.
Oracle stages: 20 exchange, 21 mass update / preclassify, 25 orphans,
26 detached advection, 27 residual compaction, 30 refill, 31 repair.
"""
import numpy as np
import pytest

import lbm_abi as abi
from harness import env_of
from lbm_config import parse_env
from oracle_ref import OracleCase

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0])
CY = np.array([0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, -1, 1])
CZ = np.array([0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
OPP = [0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12, 17, 18, 15, 16]
GAS, LIQ, INT = abi.FS_GAS, abi.FS_LIQUID, abi.FS_INTERFACE
N = 16


def box(**kw):
    """A closed 16^3 water box (solid x and y walls, flat bed below the
    grid), every cell open; states are uploaded by the tests."""
    base = dict(LBM_LX_M=N, LBM_LY_M=N, LBM_LZ_M=N, LBM_FREE_SURFACE_DEPTH_M=8, LBM_VOF_INITIAL_SHAPE="none",
                LBM_G_LB=1e-4)
    base.update({k: str(v) for k, v in kw.items()})
    cfg = parse_env(env_of("vof_192x96x160", **base))
    o = OracleCase(cfg.to_struct(), fp64=True)
    o.geometry(0.0)
    assert np.all(o.download(abi.FIELD_OBSTACLE) == 0)
    return cfg, o


def upload_state(o, ty, fill, mass=None, rho=None, u=None, f=None, ty_prev=None):
    sh = (N, N, N)
    o.upload(abi.FIELD_FS_TYPE, ty.astype(np.int32))
    o.upload(abi.FIELD_FS_TYPE_PREV, (ty if ty_prev is None else ty_prev).astype(np.int32))
    o.upload(abi.FIELD_FS_FILL, fill)
    o.upload(abi.FIELD_FS_MASS, fill * 1.0 if mass is None else mass)
    o.upload(abi.FIELD_RHO, np.ones(sh) if rho is None else rho)
    uu = (np.zeros(sh),) * 3 if u is None else u
    for fid, a in zip((abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ), uu):
        o.upload(fid, np.broadcast_to(a, sh).copy())
    if f is not None:
        o.upload(abi.FIELD_F, f)


# ---- a17: mass exchange ------------------------------------------------------
def test_exchange_is_antisymmetric_between_interface_cells():
    """Every open cell an interface cell with a random fill: each link's
    exchange is counted once with each sign, so the box total is zero."""
    cfg, o = box()
    rng = np.random.default_rng(1)
    sh = (N, N, N)
    fill = rng.uniform(0.3, 0.7, sh)
    f = W[:, None, None, None] * (1 + 0.1 * rng.uniform(-1, 1, (19,) + sh))
    upload_state(o, np.full(sh, INT), fill, mass=0.5 * np.ones(sh), f=f)
    o.stage(20)
    o.stage(21)
    dm = o.download(abi.FIELD_FS_MASS) - 0.5
    assert np.max(np.abs(dm)) > 1e-4  # the exchange is not trivially zero
    assert abs(dm.sum()) < 1e-13
    o.close()


@pytest.mark.parametrize("gas_above", [False, True])
def test_exchange_of_a_linear_population_field(gas_above):
    """f_q(x) = g_q + h_q x around one interface cell among liquid cells:
    dm = sum_q f_qbar(x) - f_q(x + c_q) = -sum_q h_q c_x,q.  A gas cell
    above removes the link q = 5 (source above) from the sum."""
    cfg, o = box()
    rng = np.random.default_rng(2)
    sh = (N, N, N)
    g = W * (1 + 0.1 * rng.uniform(-1, 1, 19))
    h = 1e-4 * rng.uniform(-1, 1, 19)
    x = np.arange(N, dtype=float)
    f = g[:, None, None, None] + h[:, None, None, None] * x[None, None, None, :]
    f = np.broadcast_to(f, (19,) + sh).copy()
    ty = np.full(sh, LIQ)
    fill = np.ones(sh)
    c = (8, 8, 8)  # (k, j, i)
    ty[c] = INT
    fill[c] = 0.5
    if gas_above:
        ty[9, 8, 8] = GAS
        fill[9, 8, 8] = 0.0
    mass = fill.copy()
    upload_state(o, ty, fill, mass=mass, f=f)
    o.stage(20)
    o.stage(21)
    dm = o.download(abi.FIELD_FS_MASS)[c] - 0.5
    exp = -np.sum(h * CX)
    if gas_above:
        xi = x[8]
        exp -= (g[6] + h[6] * xi) - (g[5] + h[5] * xi)  # the q = 5 term of the sum
    assert abs(dm - exp) < 1e-15, (dm, exp)
    o.close()


# ---- a22: orphan removal -----------------------------------------------------
def _orphan_layout():
    """An isolated wet cell O and three resolved interface receivers on its
    rays (2 along +x, 3 along +z, 2 along the (1,1,0) diagonal), each
    paired with a companion off every ray so that it is not an orphan."""
    o_ = np.array([7, 7, 7])
    rays = [(np.array([0, 0, 1]), 2), (np.array([1, 0, 0]), 3), (np.array([0, 1, 1]), 2)]  # (dk, dj, di)
    recv = [tuple(o_ + d * r) for r, d in rays]
    comp = [tuple(np.array(p) + e) for p, e in zip(recv, [(0, 1, 0), (0, 0, 1), (1, 0, 0)])]
    return tuple(o_), recv, comp, [d * d * int(np.sum(r * r)) for r, d in rays]


@pytest.mark.parametrize("room", [True, False])
def test_orphan_mass_goes_to_the_rays_by_room_over_d2(room):
    cfg, o = box()
    sh = (N, N, N)
    ty = np.full(sh, GAS)
    fill = np.zeros(sh)
    O, recv, comp, d2 = _orphan_layout()
    ty[O] = INT
    fill[O] = 0.3
    fills = [0.5, 0.6, 0.8]
    for p, q, fv in zip(recv, comp, fills):
        ty[p] = INT
        fill[p] = fv
        ty[q] = INT
        fill[q] = 0.9
    rho = np.ones(sh)
    mass = fill * rho
    if not room:  # receivers full: the topology weights fill / d^2 take over
        for p in recv:
            mass[p] = rho[p]
    upload_state(o, ty, fill, mass=mass, rho=rho)
    o.stage(25)
    m1 = o.download(abi.FIELD_FS_MASS)
    assert o.download(abi.FIELD_FS_TYPE)[O] == GAS and m1[O] == 0 and o.download(abi.FIELD_FS_FILL)[O] == 0
    assert o.counters()["free_surface_vof_orphan_count"] == 1
    w = np.array([(rho[p] - mass[p]) / d for p, d in zip(recv, d2)]) if room else np.array(
        [fill[p] / d for p, d in zip(recv, d2)])
    share = 0.3 * w / w.sum()
    for p, s in zip(recv, share):
        assert abs(m1[p] - (mass[p] + s)) < 1e-15, (p, m1[p], mass[p] + s)
    for q in comp:
        assert m1[q] == mass[q]
    assert abs(m1.sum() - mass.sum()) < 1e-14
    o.close()


# ---- a23: detached advection and residual compaction ------------------------
@pytest.mark.parametrize("u,capped", [((0.1, -0.05, 0.2), False), ((0.6, 0.0, 0.7), True)])
def test_detached_cell_advects_downstream(u, capped):
    cfg, o = box()
    sh = (N, N, N)
    ty = np.full(sh, GAS)
    fill = np.zeros(sh)
    c = (8, 8, 8)
    ty[c] = INT
    fill[c] = 0.4
    mass = fill.copy()
    upload_state(o, ty, fill, mass=mass, u=tuple(np.full(sh, v) for v in u))
    o.stage(26)
    m1 = o.download(abi.FIELD_FS_MASS)
    ux, uy, uz = u
    rates = {(0, 0, 1): max(ux, 0), (0, 0, -1): max(-ux, 0), (0, 1, 0): max(uy, 0), (0, -1, 0): max(-uy, 0),
             (1, 0, 0): max(uz, 0), (-1, 0, 0): max(-uz, 0)}
    tot = sum(rates.values())
    scale = 0.95 / tot if tot > 0.95 else 1.0
    moved = 0.0
    for (dk, dj, di), r in rates.items():
        exp = 0.4 * r * scale
        assert abs(m1[8 + dk, 8 + dj, 8 + di] - exp) < 1e-16
        moved += exp
    assert abs(m1[c] - (0.4 - moved)) < 1e-15
    if capped:
        assert abs(moved - 0.95 * 0.4) < 1e-15
    cnt = o.counters()
    assert cnt["free_surface_vof_detached_advect_count"] == 1
    assert abs(cnt["free_surface_vof_detached_advect_mass"] - moved) < 1e-15
    assert abs(m1.sum() - 0.4) < 1e-15
    o.close()


def test_residual_cell_moves_whole_mass_to_the_fastest_gas_neighbour():
    cfg, o = box()
    sh = (N, N, N)
    ty = np.full(sh, GAS)
    fill = np.zeros(sh)
    c = (8, 8, 8)
    ty[c] = INT
    fill[c] = 0.005  # below the residual threshold 0.01
    mass = fill.copy()
    upload_state(o, ty, fill, mass=mass, u=(np.full(sh, 0.1), np.zeros(sh), np.full(sh, -0.3)))
    o.stage(27)
    m1 = o.download(abi.FIELD_FS_MASS)
    assert m1[c] == 0.0
    assert m1[7, 8, 8] == 0.005  # -z: rate 0.3 beats +x: 0.1
    assert o.counters()["free_surface_vof_detached_compress_count"] == 1
    o.close()


# ---- a26: refill -------------------------------------------------------------
def _refill_state(rng):
    """A refilled cell R (gas before, interface now) with four wet donors of
    known fill, density, velocity and populations; the other neighbours are
    gas (not donors)."""
    sh = (N, N, N)
    ty = np.full(sh, GAS)
    ty_prev = np.full(sh, GAS)
    fill = np.zeros(sh)
    rho = np.ones(sh)
    u = [np.zeros(sh) for _ in range(3)]
    f = np.broadcast_to(W[:, None, None, None], (19,) + sh).copy()
    R = (8, 8, 8)
    ty[R] = INT
    fill[R] = 0.2
    donors = {1: 0.5, 3: 1.0, 7: 0.25, 17: 0.75}  # link q from R -> fill
    for q, fv in donors.items():
        p = (8 + CZ[q], 8 + CY[q], 8 + CX[q])
        ty[p] = ty_prev[p] = INT if fv < 1 else LIQ
        fill[p] = fv
        rho[p] = 1 + 0.05 * rng.uniform(-1, 1)
        for a in range(3):
            u[a][p] = 0.03 * rng.uniform(-1, 1)
        f[:, p[0], p[1], p[2]] = W * (1 + 0.05 * rng.uniform(-1, 1, 19))
    return R, donors, ty, ty_prev, fill, rho, u, f


def _feq(rho, u):
    cu = CX * u[0] + CY * u[1] + CZ * u[2]
    return rho * W * (1 + 3 * cu + 4.5 * cu**2 - 1.5 * np.dot(u, u))


@pytest.mark.parametrize("mode", ["equilibrium", "noneq", "directional"])
def test_refill_takes_the_weighted_donor_state(mode):
    cfg, o = box(LBM_FREE_SURFACE_REFILL=mode)
    rng = np.random.default_rng(6)
    R, donors, ty, ty_prev, fill, rho, u, f = _refill_state(rng)
    upload_state(o, ty, fill, rho=rho, u=tuple(u), f=f, ty_prev=ty_prev)
    o.stage(30)
    ps = {q: (8 + CZ[q], 8 + CY[q], 8 + CX[q]) for q in donors}
    wts = {q: W[q] * fv for q, fv in donors.items()}
    ws = sum(wts.values())
    rho_r = sum(wts[q] * rho[ps[q]] for q in donors) / ws
    u_r = np.array([sum(wts[q] * u[a][ps[q]] for q in donors) / ws for a in range(3)])
    got_f = o.download(abi.FIELD_F)[(slice(None),) + R]
    got_rho = o.download(abi.FIELD_RHO)[R]
    got_u = np.array([o.download(x)[R] for x in (abi.FIELD_UX, abi.FIELD_UY, abi.FIELD_UZ)])
    if mode == "equilibrium":
        exp = _feq(rho_r, u_r)
    elif mode == "noneq":
        neq = sum(wts[q] * (f[(slice(None),) + ps[q]] - _feq(rho[ps[q]], np.array([u[a][ps[q]] for a in range(3)])))
                  for q in donors) / ws
        exp = _feq(rho_r, u_r) + neq
    else:
        exp = _feq(rho_r, u_r)
        for q in range(19):  # a donor at the source x - c_q gives its own population
            src = (8 - CZ[q], 8 - CY[q], 8 - CX[q])
            if any(src == p for p in ps.values()):
                exp[q] = f[(q,) + src]
        exp = exp * (rho_r / exp.sum())
    assert np.allclose(got_f, exp, rtol=0, atol=1e-15), (mode, got_f - exp)
    if mode == "directional":
        assert abs(got_rho - rho_r) < 1e-15
        assert np.allclose(got_u, (np.stack([CX, CY, CZ]) @ exp) / exp.sum(), atol=1e-15)
    else:
        assert abs(got_rho - rho_r) < 1e-15 and np.allclose(got_u, u_r, atol=1e-16)
    assert o.counters()["free_surface_refill_count"] == 1
    o.close()


def test_refill_without_wet_donor_fill_falls_back_to_weights():
    """Donors with zero fill: the weights w_q alone average the donors."""
    cfg, o = box()
    rng = np.random.default_rng(7)
    R, donors, ty, ty_prev, fill, rho, u, f = _refill_state(rng)
    for q in donors:
        fill[8 + CZ[q], 8 + CY[q], 8 + CX[q]] = 0.0
    upload_state(o, ty, fill, rho=rho, u=tuple(u), f=f, ty_prev=ty_prev)
    o.stage(30)
    ps = {q: (8 + CZ[q], 8 + CY[q], 8 + CX[q]) for q in donors}
    ws = sum(W[q] for q in donors)
    rho_r = sum(W[q] * rho[ps[q]] for q in donors) / ws
    assert abs(o.download(abi.FIELD_RHO)[R] - rho_r) < 1e-15
    o.close()


# ---- a27: coalescence repair --------------------------------------------------
def test_repair_copies_the_newly_wet_source_populations():
    cfg, o = box()
    rng = np.random.default_rng(9)
    sh = (N, N, N)
    ty = np.full(sh, LIQ)
    ty_prev = np.full(sh, LIQ)
    fill = np.ones(sh)
    f = W[:, None, None, None] * (1 + 0.05 * rng.uniform(-1, 1, (19,) + sh))
    C = (8, 8, 8)
    newly = [1, 11, 16]  # links q whose source C - c_q was gas and is wet now
    for q in newly:
        s_ = (8 - CZ[q], 8 - CY[q], 8 - CX[q])
        ty_prev[s_] = GAS
        ty[s_] = INT
        fill[s_] = 0.5
    upload_state(o, ty, fill, f=f, ty_prev=ty_prev)
    o.stage(31)
    got = o.download(abi.FIELD_F)[(slice(None),) + C]
    exp = f[(slice(None),) + C].copy()
    for q in newly:
        exp[q] = f[(q, 8 - CZ[q], 8 - CY[q], 8 - CX[q])]
    assert np.array_equal(got, exp)
    rho = exp.sum()
    assert abs(o.download(abi.FIELD_RHO)[C] - rho) < 1e-15
    assert abs(o.download(abi.FIELD_UX)[C] - (CX @ exp) / rho) < 1e-16
    # the source cells themselves (newly wet) are not repaired
    for q in newly:
        s_ = (slice(None), 8 - CZ[q], 8 - CY[q], 8 - CX[q])
        assert np.array_equal(o.download(abi.FIELD_F)[s_], f[s_])
    assert o.counters()["free_surface_coalescence_repair_count"] >= len(newly)
    o.close()
