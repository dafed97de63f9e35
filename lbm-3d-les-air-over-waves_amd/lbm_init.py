"""Host-side initial state of a case (the "initialise" half of the drop-in).

Restates the host/NumPy parts of the reference's init sequence
(LBM_3D_LES_air_over_waves_CODEX.py:3777-3792, "CODEX"):

  * the log-law inlet profile and the randomised initial velocity of
    init_fields (CODEX:2714-2749), with NumPy's PCG64 in place of Taichi's
    RNG (the reference seeds ti.random through run_ab_case.py:261-273);
  * initialize_free_surface_height + classify_free_surface_cells +
    apply_vof_initial_disconnected_shape (CODEX:1197-1287) for VOF pools;
  * prepare_initial_density_field (CODEX:743-818) for flat pools;
  * the water branch of init_fields (CODEX:2751-2760).

New initial states (LBM_INITIAL_STATE): uniform_noise (u0 + U(-a, a) per
component) and loglaw_noise (log-law mean profile with uniform relative
perturbations, also used as the inflow plane).  All draws use
numpy.random.default_rng(seed), arrays in C order (nz, ny, nx), components
ux -> uy -> uz.

Transcendentals are evaluated in double by the C library (math.exp/log/cos/
sin) and rounded once where the value is float32, and sums run in index
order, so the library's own restatement (lbm_initial_fields in
csrc/lbm_init.cpp) reproduces every array bit for bit.

The device finishes the sequence (solid-cell wall velocities, f = feq,
VOF mass, interface layer) in lbm_init_populations.
"""
from __future__ import annotations

import math

import numpy as np

import lbm_abi as abi
from lbm_config import SolverConfig


def inlet_profile(cfg: SolverConfig) -> np.ndarray:
    """u_inlet_profile of init_fields (CODEX:2719-2738), float32 arithmetic."""
    f32 = np.float32
    B = f32(5.0)
    u_tau = f32(0.1) * f32(cfg.u_top_lb)
    nu_lbm = (f32(cfg.tau0) - f32(0.5)) / f32(3.0)
    k = np.arange(cfg.Nz, dtype=np.float32)
    y_plus = k * u_tau / nu_lbm
    y_plus = np.where(y_plus <= f32(1.0), f32(1.0), y_plus).astype(np.float32)
    tmp = (f32(1.0) / f32(cfg.kappa)) * _f32_map(math.log, y_plus) + B
    top = tmp[-1] if tmp[-1] != 0.0 else f32(1.0)
    return (tmp / top * f32(cfg.u_top_lb)).astype(np.float32)


def loglaw_profile(cfg: SolverConfig) -> np.ndarray:
    """Log-law mean profile (new): u(z) = A ln(max(z - z_b, z0) / z0) with
    z_b the bed level, z0 in cells, A set so the mean over the cells above the
    bed equals LBM_INFLOW_U_LU."""
    dx = cfg.dx
    z = (np.arange(cfg.Nz, dtype=np.float64) + 0.5) * dx
    z_rel = z - cfg.offset_m
    z0 = cfg.inflow_z0_cells * dx
    prof = np.array([math.log(v) for v in np.maximum(z_rel, z0) / z0], dtype=np.float64)
    above = z_rel > 0.0
    mean = _seq_sum(prof[above]) / int(above.sum()) if np.any(above) else 1.0
    return (cfg.inflow_u_lu * prof / mean).astype(np.float64)


def _f32_map(fn, x: np.ndarray) -> np.ndarray:
    """fn (a math-module function) of each float32 value, evaluated in
    double and rounded once to float32."""
    flat = np.asarray(x, dtype=np.float32).ravel()
    return np.array([fn(float(v)) for v in flat], dtype=np.float64).astype(np.float32).reshape(np.shape(x))


def _seq_sum(x: np.ndarray) -> float:
    """Sum in index order (numpy's pairwise sum rounds differently)."""
    s = 0.0
    for v in np.asarray(x, dtype=np.float64).ravel():
        s += float(v)
    return s


def initial_fields(cfg: SolverConfig, obstacle0: np.ndarray, bed: np.ndarray = None) -> dict:
    """Macroscopic initial fields and boundary arrays to upload.

    obstacle0 is the t=0 solid mask (1 = no open link) from the device
    geometry; it classifies closed cells as gas in water runs.  bed is the
    bed height water_h at t = 0 (default: the flat bed level)."""
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    real = np.float64 if cfg.storage == "fp64" else np.float32
    rng = np.random.default_rng(cfg.seed)
    out: dict = {}
    u_inlet = inlet_profile(cfg)
    out[abi.FIELD_U_INLET_PROFILE] = u_inlet.astype(real)
    rho = np.full((nz, ny, nx), cfg.rho0, dtype=real)
    uy = np.zeros((nz, ny, nx), dtype=real)
    uz = np.zeros((nz, ny, nx), dtype=real)
    inflow = np.zeros((nz, ny), dtype=real)
    if cfg.initial_state == "uniform_noise":
        a = cfg.init_noise_lu
        noise = rng.uniform(-a, a, size=(3, nz, ny, nx))
        ux = (cfg.init_u_lu + noise[0]).astype(real)
        uy = noise[1].astype(real)
        uz = noise[2].astype(real)
    elif cfg.initial_state == "loglaw_noise":
        prof = loglaw_profile(cfg)
        pert = rng.uniform(-1.0, 1.0, size=(nz, ny, nx))
        ux = (prof[:, None, None] * (1.0 + cfg.inflow_noise * pert)).astype(real)
        pin = rng.uniform(-1.0, 1.0, size=(nz, ny))
        inflow = (prof[:, None] * (1.0 + cfg.inflow_noise * pin)).astype(real)
    else:
        # init_fields: ux = u_inlet_profile[k] * (1 + (2U - 1) * TI)
        rand_kick = (rng.random((nz, ny, nx), dtype=np.float32) - np.float32(0.5)) * np.float32(2.0)
        ux = (u_inlet[:, None, None] * (np.float32(1.0) + rand_kick * np.float32(cfg.TI))).astype(real)
    out[abi.FIELD_INFLOW_UX] = inflow
    if cfg.water_free_surface_enabled:
        fill, ctype, fs_h = free_surface_state(cfg, obstacle0)
        if bed is None:
            bed = np.full((ny, nx), cfg.offset_m, dtype=np.float32)
        rho_init = initial_density(cfg, fs_h, bed)
        gas = ctype == abi.FS_GAS
        rho = np.where(gas, cfg.rho0, rho_init).astype(real)
        ux0, uz0 = solitary_velocity(cfg, fs_h, bed)
        ux = np.where(gas, 0.0, ux0).astype(real)
        uy = np.zeros_like(ux)
        uz = np.where(gas, 0.0, uz0).astype(real)
        out[abi.FIELD_FS_FILL] = fill.astype(real)
        out[abi.FIELD_FS_TYPE] = ctype.astype(np.int32)
        out[abi.FIELD_FS_TYPE_PREV] = ctype.astype(np.int32)
        out[abi.FIELD_FS_H] = fs_h.astype(real)
    out[abi.FIELD_RHO] = rho
    out[abi.FIELD_UX] = ux
    out[abi.FIELD_UY] = uy
    out[abi.FIELD_UZ] = uz
    return out


def _periodic_offset(cfg: SolverConfig, xrel: np.ndarray) -> np.ndarray:
    """x - x0 folded into [-Lx/2, Lx/2] on a periodic x axis, float32."""
    f32 = np.float32
    if cfg.x_boundary_mode != "periodic":
        return xrel
    lx = f32(cfg.Nx) * f32(cfg.dx)
    xrel = np.where(xrel > f32(0.5) * lx, xrel - lx, xrel).astype(np.float32)
    return np.where(xrel < f32(-0.5) * lx, xrel + lx, xrel).astype(np.float32)


def solitary_profile(cfg: SolverConfig, t: float = 0.0):
    """KdV solitary wave of amplitude A on depth h0 (CODEX:1166-1182):
    eta(x) = A sech^2(kappa (x - x0 - dir c t)), kappa = sqrt(3 A / (4 h0^3)),
    float32 at the cell centres.  Returns (sech, tanh) of the argument, so
    the height and its slope follow without a second evaluation."""
    f32 = np.float32
    dx = f32(cfg.dx)
    h0 = f32(cfg.free_surface_depth_m)
    amp = f32(cfg.solitary_amp_m)
    kappa = np.sqrt(f32(3.0) * amp / (f32(4.0) * h0 * h0 * h0)).astype(np.float32)
    center = f32(cfg.solitary_x0_m) + f32(cfg.solitary_direction) * f32(cfg.solitary_celerity_phys) * f32(t)
    xp = ((np.arange(cfg.Nx, dtype=np.float32) + f32(0.5)) * dx).astype(np.float32)
    arg = (kappa * _periodic_offset(cfg, (xp - center).astype(np.float32))).astype(np.float32)
    ep = _f32_map(math.exp, arg)
    em = _f32_map(math.exp, -arg)
    den = (ep + em).astype(np.float32)
    return (f32(2.0) / den).astype(np.float32), ((ep - em) / den).astype(np.float32)


def surface_height(cfg: SolverConfig) -> np.ndarray:
    """initialize_free_surface_height (CODEX:1197-1207): flat, the Gaussian
    hump gaussian_surface_height_at_x (CODEX:1185-1194) or the solitary wave
    at t = 0, float32."""
    f32 = np.float32
    ny, nx = cfg.Ny, cfg.Nx
    dx = f32(cfg.dx)
    if cfg.free_surface_initial_condition == "solitary":
        sech, _ = solitary_profile(cfg)
        h = (f32(cfg.free_surface_base_m) + f32(cfg.solitary_amp_m) * sech * sech).astype(np.float32)
        return np.broadcast_to(h[None, :], (ny, nx)).astype(np.float32)
    if cfg.free_surface_initial_condition != "gaussian":
        return np.full((ny, nx), f32(cfg.free_surface_base_m), dtype=np.float32)
    lx = f32(nx) * dx
    xp = ((np.arange(nx, dtype=np.float32) + f32(0.5)) * dx).astype(np.float32)
    xrel = (xp - f32(cfg.gaussian_center_x_m)).astype(np.float32)
    if cfg.x_boundary_mode == "periodic":
        xrel = np.where(xrel > f32(0.5) * lx, xrel - lx, xrel).astype(np.float32)
        xrel = np.where(xrel < f32(-0.5) * lx, xrel + lx, xrel).astype(np.float32)
    arg = (xrel / f32(cfg.gaussian_sigma_m)).astype(np.float32)
    h = (f32(cfg.free_surface_base_m) + f32(cfg.gaussian_amp_m) * _f32_map(math.exp, f32(-0.5) * arg * arg)).astype(np.float32)
    return np.broadcast_to(h[None, :], (ny, nx)).astype(np.float32)


def free_surface_state(cfg: SolverConfig, obstacle0: np.ndarray):
    """initialize_free_surface_height, classify_free_surface_cells and
    apply_vof_initial_disconnected_shape (CODEX:1197-1287) in float32."""
    f32 = np.float32
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    dx = f32(cfg.dx)
    fs_h = surface_height(cfg)
    z_bottom = np.arange(nz, dtype=np.float32)[:, None, None] * dx
    fill = ((fs_h[None, :, :] - z_bottom) / dx).astype(np.float32)
    fill = np.clip(fill, f32(0.0), f32(1.0))
    ctype = np.where(fill <= f32(1e-4), abi.FS_GAS, np.where(fill >= f32(1.0) - f32(1e-4), abi.FS_LIQUID, abi.FS_INTERFACE))
    closed = obstacle0 != 0
    fill = np.where(closed, f32(0.0), fill).astype(np.float32)
    ctype = np.where(closed, abi.FS_GAS, ctype).astype(np.int32)
    if cfg.vof_initial_shape == "block":
        def overlap(lo, hi, bmin, bmax):
            a = np.maximum(lo, f32(bmin))
            b = np.minimum(hi, f32(bmax))
            ov = np.maximum(b - a, f32(0.0))
            return (ov / (hi - lo)).astype(np.float32)

        x0 = np.arange(nx, dtype=np.float32) * dx
        y0 = np.arange(ny, dtype=np.float32) * dx
        z0 = np.arange(nz, dtype=np.float32) * dx
        fx = overlap(x0, x0 + dx, cfg.vof_block_center_x_m - 0.5 * cfg.vof_block_size_x_m,
                     cfg.vof_block_center_x_m + 0.5 * cfg.vof_block_size_x_m)
        fy = overlap(y0, y0 + dx, cfg.vof_block_center_y_m - 0.5 * cfg.vof_block_size_y_m,
                     cfg.vof_block_center_y_m + 0.5 * cfg.vof_block_size_y_m)
        fz = overlap(z0, z0 + dx, cfg.vof_block_center_z_m - 0.5 * cfg.vof_block_size_z_m,
                     cfg.vof_block_center_z_m + 0.5 * cfg.vof_block_size_z_m)
        block = (fx[None, None, :] * fy[None, :, None] * fz[:, None, None]).astype(np.float32)
        sel = (~closed) & (block > f32(1.0e-4))
        nf = np.maximum(fill, block)
        liquid = nf >= f32(1.0) - f32(1.0e-4)
        nf = np.where(liquid, f32(1.0), nf)
        nt = np.where(liquid, abi.FS_LIQUID, abi.FS_INTERFACE)
        fill = np.where(sel, nf, fill).astype(np.float32)
        ctype = np.where(sel, nt, ctype).astype(np.int32)
    return fill, ctype, fs_h


def solitary_velocity(cfg: SolverConfig, fs_h: np.ndarray, bed: np.ndarray):
    """Initial velocity of the wet cells (CODEX:2756-2800), float32: the
    current u_current everywhere; under a solitary initial surface plus the
    shallow-water velocity dir c eta / (h0 + eta), and for
    "shallow_water_divfree" the vertical velocity -z_rel du/dx that makes it
    divergence-free (z_rel = height above the bed)."""
    f32 = np.float32
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    ux = np.full((nz, ny, nx), f32(cfg.u_current_lb), dtype=np.float32)
    uz = np.zeros((nz, ny, nx), dtype=np.float32)
    if cfg.free_surface_initial_condition != "solitary" or cfg.solitary_initial_velocity_mode == "none":
        return ux, uz
    h0 = f32(cfg.free_surface_depth_m)
    c = f32(cfg.solitary_direction) * f32(cfg.solitary_celerity_phys)
    vs = f32(cfg.vel_scale)
    eta = np.maximum((fs_h.astype(np.float32) - f32(cfg.free_surface_base_m)).astype(np.float32), f32(0.0))
    u_sol = (c * eta / (h0 + eta)).astype(np.float32)
    ux = (ux + (u_sol / vs)[None, :, :]).astype(np.float32)
    if cfg.solitary_initial_velocity_mode == "shallow_water_divfree":
        sech, tanh = solitary_profile(cfg)
        eta_x = (f32(-2.0) * f32(cfg.solitary_amp_m) * f32(cfg.solitary_kappa_m_inv) * sech * sech * tanh).astype(np.float32)
        du_dx = (c * h0 * eta_x[None, :] / ((h0 + eta) * (h0 + eta))).astype(np.float32)
        zc = ((np.arange(nz, dtype=np.float32) + f32(0.5)) * f32(cfg.dx)).astype(np.float32)
        z_rel = np.maximum(zc[:, None, None] - bed.astype(np.float32)[None, :, :], f32(0.0)).astype(np.float32)
        uz = (uz + (-z_rel * du_dx[None, :, :]) / vs).astype(np.float32)
    return ux, uz


def linear_wave_head(cfg: SolverConfig, eta: np.ndarray, bed: float) -> np.ndarray:
    """Pressure head (m) of the linear potential-flow wave under the surface
    elevation eta(x) on a flat bed (CODEX:743-818, linear_wave branch): each
    horizontal mode k of eta decays with depth as cosh(k z') / cosh(k h)
    (z' = height above the bed, capped at h = max(base - bed, dx)).  Modes:
    cos(n pi x / Lx) for solid x walls (discrete cosine series), else the
    periodic Fourier modes through a direct real DFT.  Sums run in index
    order.  Returns head[k_plane, i] in float64."""
    nx, nz, dx = cfg.Nx, cfg.Nz, cfg.dx
    eta = np.asarray(eta, dtype=np.float64)
    depth = max(cfg.free_surface_base_m - bed, dx)
    x = (np.arange(nx, dtype=np.float64) + 0.5) * dx
    z = (np.arange(nz, dtype=np.float64) + 0.5) * dx

    def decay(k, zr):
        # cosh(k zr) / cosh(k depth) without overflow: (e^{k(zr-h)} + e^{-k(zr+h)}) / (1 + e^{-2kh})
        out = np.ones_like(k)
        for m, km in enumerate(k):
            if km > 0.0:
                kh, kz = km * depth, km * zr
                out[m] = (math.exp(kz - kh) + math.exp(-kz - kh)) / (1.0 + math.exp(-2.0 * kh))
        return out

    head = np.empty((nz, nx), dtype=np.float64)
    if cfg.x_boundary_mode == "solid":
        k = np.arange(nx, dtype=np.float64) * math.pi / (nx * dx)
        basis = np.array([[math.cos(km * xi) for xi in x] for km in k])   # basis[n, i] = cos(k_n x_i)
        amp = np.zeros(nx)
        for i in range(nx):
            amp += basis[:, i] * eta[i]
        amp = (2.0 / nx) * amp
        amp[0] = _seq_sum(eta) / nx
        for kk, zc in enumerate(z):
            zr = min(max(zc - bed, 0.0), depth)
            w = amp * decay(k, zr)
            s = np.zeros(nx)
            for m in range(nx):
                s += w[m] * basis[m]
            head[kk] = s
        return head
    nm = nx // 2 + 1
    ang = [[2.0 * math.pi * ((m * i) % nx) / nx for i in range(nx)] for m in range(nm)]
    cs = np.array([[math.cos(a) for a in row] for row in ang])
    sn = np.array([[math.sin(a) for a in row] for row in ang])
    re = np.zeros(nm)
    im = np.zeros(nm)
    for i in range(nx):
        re += eta[i] * cs[:, i]
        im -= eta[i] * sn[:, i]
    k = 2.0 * math.pi * np.arange(nm, dtype=np.float64) / (nx * dx)
    nyq = nx // 2 if nx % 2 == 0 else -1
    for kk, zc in enumerate(z):
        zr = min(max(zc - bed, 0.0), depth)
        d = decay(k, zr)
        s = np.full(nx, re[0] * d[0])
        for m in range(1, nm):
            if m == nyq:
                s += d[m] * re[m] * cs[m]
            else:
                s += 2.0 * (d[m] * (re[m] * cs[m] - im[m] * sn[m]))
        head[kk] = s / nx
    return head


def initial_density(cfg: SolverConfig, fs_h: np.ndarray, bed: np.ndarray = None) -> np.ndarray:
    """prepare_initial_density_field (CODEX:743-818): the linear-wave
    pressure under a Gaussian surface on a flat bed when asked for, else the
    hydrostatic head."""
    nz, ny, nx = cfg.Nz, cfg.Ny, cfg.Nx
    if bed is None:
        bed = np.full((ny, nx), cfg.offset_m, dtype=np.float32)
    bed64 = bed.astype(np.float64)
    flat = np.allclose(bed64, bed64[0, 0], atol=1.0e-12, rtol=0.0)
    if cfg.initial_pressure_mode == "linear_wave" and cfg.free_surface_initial_condition == "gaussian" and flat:
        rows = fs_h.astype(np.float64)
        eta = np.zeros(nx)
        for j in range(ny):
            eta += rows[j]
        eta = eta / ny - cfg.free_surface_base_m
        head = linear_wave_head(cfg, eta, float(bed64[0, 0]))
        if cfg.use_hydrostatic_balanced_pressure != 1:
            z = (np.arange(nz, dtype=np.float64) + 0.5) * cfg.dx
            head = np.maximum((cfg.free_surface_base_m - z)[:, None] + head, 0.0)
        line = (cfg.rho0 + cfg.g_lb * head / cfg.dx / (1.0 / 3.0)).astype(np.float32)
        return np.ascontiguousarray(np.broadcast_to(line[:, None, :], (nz, ny, nx)))
    rho = np.empty((nz, ny, nx), dtype=np.float32)
    h = fs_h.astype(np.float64)
    z = (np.arange(nz, dtype=np.float64) + 0.5) * cfg.dx
    for kk, zc in enumerate(z):
        if cfg.use_component_balanced_pressure == 1 or cfg.use_hydrostatic_balanced_pressure == 1:
            head = h - cfg.free_surface_base_m
        else:
            head = np.maximum(h - zc, 0.0)
        rho[kk] = np.asarray(cfg.rho0 + cfg.g_lb * head / cfg.dx / (1.0 / 3.0), dtype=np.float32)
    return rho
