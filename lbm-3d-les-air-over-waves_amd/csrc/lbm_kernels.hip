// lbm_kernels.hip -- gfx950 kernels of the D3Q19 step and their launchers.
//
// Two execution paths compute the same step:
//  * staged: one kernel per stage of the reference step loop
//    (LBM_3D_LES_air_over_waves_CODEX.py:3873-3925), whole fields in HBM,
//    used for water/VOF runs and as the cross-check of the fused path;
//  * fused: air runs.  The state between steps is the post-collision
//    population set f*; a moments pass streams it (with every boundary
//    rule), conditions the wall links of the next step and writes the
//    velocity; a collision pass repeats the pull, evaluates the LES
//    viscosity from the neighbours' velocities, collides and writes f* of
//    the next step.  With FP32 storage both passes run in one launch whose
//    tiles wait on per-tile flags (the wavefront schedule).
//
// Arithmetic follows the CPU restatement operation for operation (see
// lbm_device.h), so results agree bit for bit in FP32 and FP64 storage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "lbm_device.h"
#include "lbm_internal.h"

namespace lbm {

#define HIP_OK(expr)                                                                                    \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static constexpr int NT = 256;  // threads per block of the one-cell-per-thread kernels

template <class T>
static Consts<T> make_consts(const lbm_config_t& c, int omega_band_top) {
    Consts<T> k{};
    k.nx = c.nx;
    k.ny = c.ny;
    k.nz = c.nz;
    fast_div_magic((uint32_t)c.nx, k.div_nx_m, k.div_nx_s);
    fast_div_magic((uint32_t)c.ny, k.div_ny_m, k.div_ny_s);
    k.fs_water = c.fs_enabled != 0;
    k.air_walls = c.physics_mode == LBM_PHYS_AIR;
    k.collision_reg = c.collision == LBM_COLL_REGULARIZED;
    k.use_les = c.use_les;
    k.x_boundary = c.x_boundary;
    k.y_boundary = c.y_boundary;
    k.top_boundary = c.top_boundary;
    k.bed_profile = c.bed_profile;
    k.bed_moving = c.bed_profile == LBM_BED_MOVING_SINE || c.bed_profile == LBM_BED_MOVING_SINE1;
    k.include_bed = c.include_bed;
    k.include_cube = c.include_cube;
    k.phi_fraction = c.phi_fraction;
    k.phi_flags = c.phi_flags;
    for (int a = 0; a < 3; a++) {
        k.cube_half[a] = (T)c.cube_half[a];
        k.cube_center[a] = (T)c.cube_center[a];
    }
    for (int a = 0; a < 4; a++) k.cube_rot[a] = (T)c.cube_rot[a];
    for (int a = 0; a < 5; a++) k.solitary[a] = (T)c.solitary[a];
    k.wall_link = c.wall_velocity_link;
    k.top_drive = c.enable_top_drive;
    k.x_fringe = c.enable_x_fringe;
    k.top_sponge = c.enable_top_sponge;
    k.fs_vof = c.fs_enabled && c.fs_tracking == LBM_FST_VOF;
    k.fs_mode = c.fs_mode;
    k.pf = c.pressure_formulation;
    k.fs_boundary_hydro = c.fs_boundary_mode == LBM_FSB_HYDROSTATIC;
    k.refill_mode = c.refill_mode;
    k.link_aperture = c.vof_link_aperture;
    k.thin_gap_bridge = c.vof_thin_gap_bridge;
    k.omega_band_top = omega_band_top;
    k.seed = c.seed;
    k.dx = (T)c.dx;
    k.rho0 = (T)c.rho0;
    k.tau0 = (T)c.tau0;
    k.nu0 = (T)c.nu0;
    T csd = (T)c.cs_les * (T)c.delta_les;
    k.cs2_les = csd * csd;
    k.nu_t_max = (T)c.nu_t_max;
    k.kappa = (T)c.kappa;
    k.z0_phys = (T)c.z0_phys;
    k.u_top_lb = (T)c.u_top_lb;
    k.u_current_lb = (T)c.u_current_lb;
    k.g_lb = (T)c.g_lb;
    k.g_body_lb = (T)c.g_body_lb;
    k.c_backscatter = (T)c.c_backscatter;
    k.free_surface_base_m = (T)c.free_surface_base_m;
    k.bridge_strength = (T)c.vof_thin_gap_bridge_strength;
    k.offset_m = (T)c.offset_m;
    k.bed_amp_m = (T)c.bed_amp_m;
    k.bed_wavelength_m = (T)c.bed_wavelength_m;
    k.wave_speed = (T)c.wave_speed;
    k.wave_period = (T)c.wave_period;
    k.depth_swl_m = (T)c.depth_swl_m;
    k.vel_scale = (T)c.vel_scale;
    k.phase_speed = (T)c.bed_phase_speed_mps;
    k.dt_phys = (T)c.dt_phys;
    {
        int e = 0;
        double m = frexp((double)c.dx, &e);
        k.dx_pow2 = m == 0.5 ? 1 : 0;
        k.inv_dx = k.dx_pow2 ? (T)ldexp(1.0, 1 - e) : T(1) / k.dx;
    }
    float sc = 1.0f;
    for (int i = 0; i < c.storage_shift; i++) sc *= 2.0f;
    k.store.scale = sc;
    k.store.inv_scale = 1.0f / sc;
    k.store.stochastic = c.storage_rounding == LBM_ROUND_STOCHASTIC ? 1 : 0;
    {
        // s_q = feq_q(1, u_ref) in float, the order of the equilibrium (CODEX:3326-3333)
        const float ux = c.storage_u_ref[0], uy = c.storage_u_ref[1], uz = c.storage_u_ref[2];
        const float u2 = ux * ux + uy * uy + uz * uz;
        for (int q = 0; q < 19; q++) {
            float cu = 3.0f * ((float)cxq(q) * ux + (float)cyq(q) * uy + (float)czq(q) * uz);
            k.store.shift[q] = wq<float>(q) * (1.0f + cu + 0.5f * cu * cu - 1.5f * u2);
        }
    }
    k.store_step = 0;
    return k;
}

template <class T>
static Slot<T> make_slot(DevBufs& b, int s) {
    Slot<T> sl;
    sl.h = (const T*)b.h[s];
    sl.uw = (const T*)b.uw[s];
    sl.vw = (const T*)b.vw[s];
    sl.ww = (const T*)b.ww[s];
    sl.eta9 = (const T*)b.eta9[s];
    sl.denom = (const T*)b.denom[s];
    sl.kband = b.kband[s];
    sl.uw9 = (const T*)b.uw9[s];
    sl.ksolid = b.ksolid[s];
    sl.fsol = (const T*)b.fsol[s];
    sl.ibb = (const T*)b.ibb[s];
    sl.phi = (const T*)b.phi;
    return sl;
}

static inline unsigned nblocks(size_t n, int nt = NT) { return (unsigned)((n + nt - 1) / nt); }

template <class T>
__device__ __forceinline__ void cell_ijk(const Consts<T>& c, size_t id, int& i, int& j, int& k) {
    const uint32_t u = (uint32_t)id;
    const uint32_t r = fast_div(u, c.div_nx_m, c.div_nx_s);
    const uint32_t kk = fast_div(r, c.div_ny_m, c.div_ny_s);
    i = (int)(u - r * (uint32_t)c.nx);
    j = (int)(r - kk * (uint32_t)c.ny);
    k = (int)kk;
}

// Deterministic FP64 reductions in two passes.  Pass 1: block b sums the
// cells [b * RED_CHUNK, (b + 1) * RED_CHUNK), thread t the cells
// t, t + RED_THREADS, ... of that chunk in order, then a fixed tree over the
// block; pass 2: one block sums the partials the same way.  The grouping
// depends only on the grid size, so a result does not depend on timing or
// on what else runs on the device.
constexpr int RED_THREADS = 256;
constexpr int RED_PER_THREAD = 16;
constexpr size_t RED_CHUNK = (size_t)RED_THREADS * RED_PER_THREAD;

// fixed-shape block tree: xor butterfly inside each wave, then the four
// wave results in order (sum, or max when MAX)
template <bool MAX>
__device__ __forceinline__ double block_tree(double v) {
    __shared__ double wsum[RED_THREADS / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(v, off);
        v = MAX ? fmax(v, o) : v + o;
    }
    const int w = (int)threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) wsum[w] = v;
    __syncthreads();
    double r = wsum[0];
#pragma unroll
    for (int i = 1; i < RED_THREADS / 64; i++) r = MAX ? fmax(r, wsum[i]) : r + wsum[i];
    __syncthreads();
    return r;
}

// ===========================================================================
// 2-D bed kernels
// ===========================================================================

// update_wave_bed_and_velocities (CODEX:1082-1162), one column; moving_sine1
// is the single-harmonic bed of the new configuration
template <class T>
struct BedEta {
    T eta1, eta2, eta, phase, phase2, amp_now;
};
template <class T>
__device__ __forceinline__ BedEta<T> bed_eta(const Consts<T>& c, T t, int i) {
    BedEta<T> e{T(0), T(0), T(0), T(0), T(0), T(0)};
    const T dx = c.dx;
    const T kw = T(2.0) * T(3.141592653589793) / c.bed_wavelength_m;
    const T cs = c.wave_speed;
    const T ramp = T(2.0) * c.wave_period;
    if (c.bed_profile == LBM_BED_MOVING_SINE)
        e.amp_now = t < ramp ? c.bed_amp_m * smoothstep01(t / ramp) : c.bed_amp_m;
    const T x0 = (T)i * dx;
    if (c.bed_profile == LBM_BED_MOVING_SINE) {
        T x_phase = x0 - cs * t;
        T x_phase2 = x0 - T(0.5) * cs * t;
        e.phase = T(0.5) * kw * x_phase;
        e.phase2 = kw * x_phase2;
        e.eta1 = e.amp_now * (T)det_sincos((double)e.phase, 0);
        e.eta2 = e.amp_now * (T)det_sincos((double)e.phase2, 0);
        e.eta = e.eta1 + e.eta2;
    } else if (c.bed_profile == LBM_BED_MOVING_SINE1) {
        e.phase = kw * (x0 - c.phase_speed * t);
        e.eta1 = c.bed_amp_m * (T)det_sincos((double)e.phase, 0);
        e.eta = e.eta1;
    }
    return e;
}
template <class T>
__device__ __forceinline__ T bed_height(const Consts<T>& c, const BedEta<T>& e, int i, const T* piecewise) {
    T bed_h = c.offset_m + e.eta;
    if (c.bed_profile == LBM_BED_PIECEWISE) bed_h = piecewise[i];
    return bed_h;
}
// bed height and wall velocity of column (i, .) at time t; hl: the height
// of the previous time level
template <class T>
__device__ __forceinline__ void bed_column(const Consts<T>& c, T t, int i, T hl, const T* piecewise, T& h_out, T& u_out, T& v_out,
                                           T& w_out) {
    const BedEta<T> e = bed_eta(c, t, i);
    const T kw = T(2.0) * T(3.141592653589793) / c.bed_wavelength_m;
    const T cs = c.wave_speed;
    const T bed_h = bed_height(c, e, i, piecewise);
    h_out = bed_h;
    T u_s = T(0), v_s = T(0), w_s = T(0);
    T h0 = c.depth_swl_m;
    if (c.bed_profile == LBM_BED_MOVING_SINE) {
        T eta_x1 = e.amp_now * T(0.5) * kw * (T)det_sincos((double)e.phase, 1);
        T eta_x2 = e.amp_now * kw * (T)det_sincos((double)e.phase2, 1);
        T eta_x = eta_x1 + eta_x2;
        T eta_y = T(0);
        T eta_t = (bed_h - hl) / c.dt_phys;
        u_s = cs * e.eta1 / (h0 + e.eta1) + T(0.5) * cs * e.eta2 / (h0 + e.eta2);
        w_s = eta_t + u_s * eta_x + v_s * eta_y;
    } else if (c.bed_profile == LBM_BED_MOVING_SINE1) {
        T eta_x = c.bed_amp_m * kw * (T)det_sincos((double)e.phase, 1);
        T eta_t = (bed_h - hl) / c.dt_phys;
        u_s = c.phase_speed * e.eta1 / (h0 + e.eta1);
        w_s = eta_t + u_s * eta_x;
    }
    u_out = u_s / c.vel_scale;
    v_out = v_s / c.vel_scale;
    w_out = w_s / c.vel_scale;
}

template <class T>
__global__ __launch_bounds__(NT) void k_bed2d_a(Consts<T> c, T t, T* h, T* uw, T* vw, T* ww, const T* h_last, const T* piecewise) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (size_t)c.nx * c.ny) return;
    const int i = (int)(id % (size_t)c.nx);
    bed_column(c, t, i, h_last[id], piecewise, h[id], uw[id], vw[id], ww[id]);
}

// BED_BATCH time levels of a moving bed in one launch (blockIdx.y = level).
// The previous height of level 0 is the stored one; level b > 0 recomputes
// the height of level b - 1 from its time with the same operations, so it
// equals the value that level b - 1 stores.
template <class T>
struct BedBatch {
    T t[BED_BATCH], tprev[BED_BATCH];
    T* h[BED_BATCH];
    T* uw[BED_BATCH];
    T* vw[BED_BATCH];
    T* ww[BED_BATCH];
    T* eta9[BED_BATCH];
    T* denom[BED_BATCH];
    int* kband[BED_BATCH];
    T* uw9[BED_BATCH];
    int* ksolid[BED_BATCH];
    T* fsol[BED_BATCH];
    T* ibb[BED_BATCH];
    const T* h_last0;
};
template <class T>
__global__ __launch_bounds__(NT) void k_bed2d_a_batch(Consts<T> c, BedBatch<T> a, const T* piecewise) {
    const size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (size_t)c.nx * c.ny) return;
    const int b = (int)blockIdx.y;
    const int i = (int)(id % (size_t)c.nx);
    const T hl = b == 0 ? a.h_last0[id] : bed_height(c, bed_eta(c, a.tprev[b], i), i, piecewise);
    bed_column(c, a.t[b], i, hl, piecewise, a.h[b][id], a.uw[b][id], a.vw[b][id], a.ww[b][id]);
}

// per-column link-endpoint bed heights (CODEX:2536-2569 at the 9 (cx,cy)
// endpoints), phi slope normaliser (CODEX:2451-2469), the highest k with a
// cut or closed link (cells above it are fully open: frac = 1 everywhere),
// the highest k whose links are all closed, and the per-link wall terms.
// A block holds 64 columns and 9 waves: wave e samples endpoint e of every
// column (4 fields), then the waves share the per-link terms and wave 0
// derives the per-column bounds.
constexpr int BED_COLS = 64;
// one wave per endpoint: the per-link loop reads e from the first lane
static_assert(BED_COLS == 64, "k_bed2d_b needs one 64-lane wave per endpoint");
template <class T>
__device__ __forceinline__ void bed2d_b_block(const Consts<T>& c, const T* h, T* eta9, T* denom, int* kband, const T* uw,
                                              const T* vw, const T* ww, T* uw9, int* ksolid, T* fsol, T* ibb, unsigned blk) {
    __shared__ T se[9][BED_COLS];
    __shared__ T sw[27][BED_COLS];
    const int lane = (int)(threadIdx.x % BED_COLS), e = (int)(threadIdx.x / BED_COLS);
    const size_t plane = (size_t)c.nx * c.ny;
    const size_t id = (size_t)blk * BED_COLS + lane;
    const bool live = id < plane;
    const int i = live ? (int)(id % (size_t)c.nx) : 0, j = live ? (int)(id / (size_t)c.nx) : 0;
    const T dx = c.dx;
    if (live) {
        const T xc = ((T)i + T(0.5)) * dx, yc = ((T)j + T(0.5)) * dx;
        const int a = e / 3 - 1, bb = e % 3 - 1;  // e = (a + 1) * 3 + (bb + 1)
        const Axis2d<T> sx = sample_axis(xc + T(0.5) * dx * (T)a, c.nx, dx);
        const Axis2d<T> sy = sample_axis(yc + T(0.5) * dx * (T)bb, c.ny, dx);
        const T ev = sample_at(h, sx, sy, c.nx);
        se[e][lane] = ev;
        eta9[(size_t)e * plane + id] = ev;
        // wall velocity at the link endpoint (CODEX:937-988)
        const T u = sample_at(uw, sx, sy, c.nx), v = sample_at(vw, sx, sy, c.nx), w = sample_at(ww, sx, sy, c.nx);
        sw[e][lane] = u;
        sw[9 + e][lane] = v;
        sw[18 + e][lane] = w;
        uw9[(size_t)e * plane + id] = u;
        uw9[(size_t)(9 + e) * plane + id] = v;
        uw9[(size_t)(18 + e) * plane + id] = w;
    }
    __syncthreads();
    if (!live) return;
    // per-link wall terms of this column, independent of k: the value the
    // conditioning of update_bed_populations_and_reinit (CODEX:2894-2940)
    // gives a closed link, feq(rho0, wall velocity of the link), and the
    // moving-wall term of the IBB (CODEX:3674-3712); x / y wall links have
    // wall velocity 0.  Wave e takes links e, e + 9, e + 18 (e is uniform over
    // a wave, so each link's code is compiled with its q fixed).
    const int ew = __builtin_amdgcn_readfirstlane(e);
#pragma unroll
    for (int q = 0; q < 19; q++) {
        if (q % 9 != ew) continue;
        int wt = c.include_bed ? LBM_WALL_BED : LBM_WALL_NONE;
        if (q > 0) wt = link_wt(c, link_on_wall(c, i, j, q), i, j, q);
        T ux = T(0), uy = T(0), uz = T(0);
        if (wt == LBM_WALL_BED) {
            if (c.wall_link) {
                const int ep = c9q(q);
                ux = sw[ep][lane];
                uy = sw[9 + ep][lane];
                uz = sw[18 + ep][lane];
            } else {
                ux = uw[id];
                uy = vw[id];
                uz = ww[id];
            }
        }
        fsol[(size_t)q * plane + id] = feq(c.rho0, ux, uy, uz, q);
        if (q > 0) {
            const int qb = oppq(q);
            T cuw = (T)cxq(qb) * ux + (T)cyq(qb) * uy + (T)czq(qb) * uz;
            ibb[(size_t)(q - 1) * plane + id] = T(6.0) * wq<T>(qb) * c.rho0 * cuw;
        }
    }
    if (e == 1) {
        int ip = wrap1(i + 1, c.nx), im = wrap1(i - 1, c.nx);
        int jp = wrap1(j + 1, c.ny), jm = wrap1(j - 1, c.ny);
        T dh_dx = (h[(size_t)j * c.nx + ip] - h[(size_t)j * c.nx + im]) * (T(0.5) / dx);
        if (c.x_boundary != LBM_XB_PERIODIC) {
            if (i == 0)
                dh_dx = (h[(size_t)j * c.nx + i + 1] - h[(size_t)j * c.nx + i]) / dx;
            else if (i == c.nx - 1)
                dh_dx = (h[(size_t)j * c.nx + i] - h[(size_t)j * c.nx + i - 1]) / dx;
        }
        T dh_dy = (h[(size_t)jp * c.nx + i] - h[(size_t)jm * c.nx + i]) * (T(0.5) / dx);
        T d = sqrt(T(1.0) + dh_dx * dh_dx + dh_dy * dh_dy);
        if (d < T(1e-6)) d = T(1.0);
        denom[id] = d;
    }
    if (e != 0) return;
    T e9[9];
    T hmax = T(-1e30);
#pragma unroll
    for (int p = 0; p < 9; p++) {
        e9[p] = se[p][lane];
        hmax = e9[p] > hmax ? e9[p] : hmax;
    }
    // a link of cell k has frac < 1 only when zq - eta < dx/2 with
    // zq >= k dx, i.e. k < eta/dx + 1/2; one extra cell of margin
    int kb = -1;
    if (c.include_bed) {
        double kk = ceil((double)hmax / (double)dx + 0.5) + 1.0;
        kb = kk < -1.0 ? -1 : (kk > (double)c.nz ? c.nz : (int)kk);
    }
    bool wall_col = (c.x_boundary == LBM_XB_SOLID && (i == 0 || i == c.nx - 1)) ||
                    (c.y_boundary == LBM_YB_SOLID && (j == 0 || j == c.ny - 1));
    if (wall_col) kb = c.nz;
    kband[id] = kb;
    // highest k whose 18 links all end below the bed (gap < 0 with the
    // arithmetic of link_geo).  Closedness is monotone in k, so the walk
    // starts one plane under the real-number bound when that plane is
    // closed (else at 0) and gives the same k as a walk from 0.
    auto closed = [&](int k) {
        T zc = ((T)k + T(0.5)) * dx;
        bool cl = true;
#pragma unroll
        for (int q = 1; q < 19; q++) {
            T zq = zc + T(0.5) * dx * (T)czq(q);
            T gap = zq - e9[c9q(q)];
            if (!(gap < T(0))) cl = false;
        }
        return cl;
    };
    int ks = -1;
    if (c.include_bed) {
        double m = 1e300;
#pragma unroll
        for (int q = 1; q < 19; q++) {
            const double v = (double)e9[c9q(q)] - 0.5 * (double)dx * (double)czq(q);
            m = v < m ? v : m;
        }
        const double kr = m / (double)dx - 0.5;  // closed for real k < kr
        int k0 = kr > 2.0 ? (int)fmin(ceil(kr) - 2.0, (double)(c.nz - 1)) : 0;
        if (k0 > 0 && !closed(k0)) k0 = 0;
        for (int k = k0; k < c.nz; k++) {
            if (!closed(k)) break;
            ks = k;
        }
    }
    ksolid[id] = ks;
}

template <class T>
__global__ __launch_bounds__(9 * BED_COLS) void k_bed2d_b(Consts<T> c, const T* h, T* eta9, T* denom, int* kband, const T* uw,
                                                          const T* vw, const T* ww, T* uw9, int* ksolid, T* fsol, T* ibb) {
    bed2d_b_block(c, h, eta9, denom, kband, uw, vw, ww, uw9, ksolid, fsol, ibb, blockIdx.x);
}
template <class T>
__global__ __launch_bounds__(9 * BED_COLS) void k_bed2d_b_batch(Consts<T> c, BedBatch<T> a) {
    const int b = (int)blockIdx.y;
    bed2d_b_block(c, (const T*)a.h[b], a.eta9[b], a.denom[b], a.kband[b], (const T*)a.uw[b], (const T*)a.vw[b],
                  (const T*)a.ww[b], a.uw9[b], a.ksolid[b], a.fsol[b], a.ibb[b], blockIdx.x);
}

// the bed batches of several cases in one launch each (the ensemble
// server's batches): case k's blocks start at a0 (k_bed2d_a_multi) and b0
// (k_bed2d_b_multi); within a case, block x of level l is block
// l * (blocks per level) + x, as blockIdx.(x, y) of the one-case launch
template <class T>
struct BedCase {
    Consts<T> c;
    BedBatch<T> a;
    const T* piecewise;
    int nb;      // levels
    int a0, b0;  // first block of the case in each launch
    int pad;
};
template <class T>
__device__ __forceinline__ int bed_case_of(const BedCase<T>* tab, int ncases, int t, bool second) {
    int lo = 0, hi = ncases - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((second ? tab[mid].b0 : tab[mid].a0) <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}
template <class T>
__global__ __launch_bounds__(64) void k_bed2d_a_multi(const BedCase<T>* __restrict__ tab, int ncases) {
    const int k = bed_case_of(tab, ncases, (int)blockIdx.x, false);
    const Consts<T> c = tab[k].c;
    const int per = (int)(((size_t)c.nx * c.ny + 63) / 64);
    const int local = (int)blockIdx.x - tab[k].a0;
    const int b = local / per;
    const size_t id = (size_t)(local - b * per) * 64 + threadIdx.x;
    if (id >= (size_t)c.nx * c.ny) return;
    const BedBatch<T>& a = tab[k].a;
    const int i = (int)(id % (size_t)c.nx);
    const T hl = b == 0 ? a.h_last0[id] : bed_height(c, bed_eta(c, a.tprev[b], i), i, tab[k].piecewise);
    bed_column(c, a.t[b], i, hl, tab[k].piecewise, a.h[b][id], a.uw[b][id], a.vw[b][id], a.ww[b][id]);
}
template <class T>
__global__ __launch_bounds__(9 * BED_COLS) void k_bed2d_b_multi(const BedCase<T>* __restrict__ tab, int ncases) {
    const int k = bed_case_of(tab, ncases, (int)blockIdx.x, true);
    const Consts<T> c = tab[k].c;
    const int per = (int)(((size_t)c.nx * c.ny + BED_COLS - 1) / BED_COLS);
    const int local = (int)blockIdx.x - tab[k].b0;
    const int b = local / per;
    const BedBatch<T>& a = tab[k].a;
    bed2d_b_block(c, (const T*)a.h[b], a.eta9[b], a.denom[b], a.kband[b], (const T*)a.uw[b], (const T*)a.vw[b], (const T*)a.ww[b],
                  a.uw9[b], a.ksolid[b], a.fsol[b], a.ibb[b], (unsigned)(local - b * per));
}

// ===========================================================================
// geometry materialisation (downloads) and per-cell flags
// ===========================================================================
template <class T>
__global__ __launch_bounds__(NT) void k_geometry_fields(Consts<T> c, Slot<T> s, int32_t* lo, T* lf, int32_t* wt, int32_t* near, T* phi) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    CellGeo<T> g;
    cell_geometry(c, s, i, j, k, g, true);
    for (int q = 0; q < 19; q++) {
        lo[(size_t)q * n + id] = (g.open_mask >> q) & 1u;
        lf[(size_t)q * n + id] = g.lf[q];
        wt[(size_t)q * n + id] = link_wall_type(c, g, i, j, q);
    }
    near[id] = g.near;
    phi[id] = cell_phi(c, s, i, j, k);
}

template <class T>
__global__ __launch_bounds__(NT) void k_cellflags(Consts<T> c, Slot<T> s, uint8_t* flags) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    CellGeo<T> g;
    cell_geometry(c, s, i, j, k, g);
    flags[id] = (uint8_t)((g.open_mask & 1u) | (g.near ? 2u : 0u));
}

// gravity of one cell for the Guo force (CODEX:646-651 with
// update_pressure_reference_fields CODEX:710-740 evaluated in place)
template <class T>
__device__ __forceinline__ T gravity_fraction(const Consts<T>& c, int k, bool open0, int type) {
    if (!(c.fs_water && open0 && type != LBM_FS_GAS)) return T(0);
    if (c.pf == LBM_PF_HYDROSTATIC_BALANCED) return T(0);
    if (c.pf == LBM_PF_VOF_COMPONENT_BALANCED) {
        T zc = ((T)k + T(0.5)) * c.dx;
        return zc < c.free_surface_base_m ? T(0) : T(1);
    }
    return T(1);
}
template <class T>
__device__ __forceinline__ T local_g(const Consts<T>& c, int k, bool open0, int type) {
    if (c.fs_water && c.pf == LBM_PF_VOF_COMPONENT_BALANCED) return c.g_lb * gravity_fraction(c, k, open0, type);
    return c.g_body_lb;
}

// ===========================================================================
// initialisation: init_fields populations (CODEX:2711-2818)
// ===========================================================================
template <class S, class T>
__global__ __launch_bounds__(NT) void k_init_populations(Consts<T> c, Slot<T> s, S* f, T* rho, T* ux, T* uy, T* uz, int* obst, const int* type) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    CellGeo<T> g;
    cell_geometry(c, s, i, j, k, g);
    bool open0 = g.open_mask & 1u;
    obst[id] = open0 ? 0 : 1;
    if (!open0) {
        T a = T(0), bq = T(0), e = T(0);
        if (link_wall_type(c, g, i, j, 0) == LBM_WALL_BED) {
            size_t col = (size_t)j * c.nx + i;
            a = s.uw[col];
            bq = s.vw[col];
            e = s.ww[col];
        }
        ux[id] = a;
        uy[id] = bq;
        uz[id] = e;
        rho[id] = c.rho0;
    }
    T u = ux[id], v = uy[id], w = uz[id];
    int ty = c.fs_water ? type[id] : LBM_FS_LIQUID;
    if (c.fs_water && open0 && ty != LBM_FS_GAS) w += T(0.5) * local_g(c, k, open0, ty);
    T r = rho[id];
    T u2 = u * u + v * v + w * w;
    const uint32_t dbase = dither_base(c.store_step, (uint32_t)id);
#pragma unroll
    for (int q = 0; q < 19; q++) {
        T cu = T(3.0) * ((T)cxq(q) * u + (T)cyq(q) * v + (T)czq(q) * w);
        T fv = r * wq<T>(q) * (T(1.0) + cu + T(0.5) * cu * cu - T(1.5) * u2);
        Store<S, T>::store(f, (size_t)q * n + id, q, fv, c.store, dbase);
    }
}

// ===========================================================================
// staged path: one kernel per reference stage
// ===========================================================================

// update_bed_populations_and_reinit (CODEX:2878-2940)
// S != T (16-bit storage): a population the conditioning changed is stored
// (rounded, dither of step ^ SALT_COND); the others keep their stored bits
template <class S, class T>
__global__ __launch_bounds__(NT) void k_cond(Consts<T> c, Slot<T> s, S* f, int* obst, int* obst_prev, uint32_t step) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    CellGeo<T> g;
    cell_geometry(c, s, i, j, k, g);
    obst_prev[id] = obst[id];
    obst[id] = (g.open_mask & 1u) ? 0 : 1;
    if ((g.open_mask & 0x7ffffu) == 0x7ffffu) return;  // no closed link
    if (!c.air_walls && !c.bed_moving) return;          // water: only a moving bed conditions its links
    T fl[19], f0[19];
#pragma unroll
    for (int q = 0; q < 19; q++) f0[q] = fl[q] = Store<S, T>::load(f, (size_t)q * n + id, q, c.store);
    condition_cell(c, s, g, i, j, fl);
    if constexpr (std::is_same<S, T>::value) {
#pragma unroll
        for (int q = 0; q < 19; q++) f[(size_t)q * n + id] = fl[q];
    } else {
        const uint32_t base = dither_base(step ^ SALT_COND, (uint32_t)id);
#pragma unroll
        for (int q = 0; q < 19; q++)
            if (__float_as_uint(fl[q]) != __float_as_uint(f0[q])) Store<S, T>::store(f, (size_t)q * n + id, q, fl[q], c.store, base);
    }
}

// macro_step (CODEX:2943-3008)
// rho and u of one cell from its populations (macro_step, CODEX:2943-3008)
template <class T>
__device__ __forceinline__ void macro_cell(const Consts<T>& c, int k, const T fl[19], int ty, bool open0, T& r, T& u, T& v, T& w) {
    r = T(0);
#pragma unroll
    for (int q = 0; q < 19; q++) r += fl[q];
    T jx = T(0), jy = T(0), jz = T(0);
#pragma unroll
    for (int q = 0; q < 19; q++) {
        if (cxq(q)) jx += fl[q] * (T)cxq(q);
        if (cyq(q)) jy += fl[q] * (T)cyq(q);
        if (czq(q)) jz += fl[q] * (T)czq(q);
    }
    if (c.fs_water) {
        if (ty == LBM_FS_GAS) {
            r = c.rho0;
            u = v = w = T(0);
        } else {
            T force_z = -r * local_g(c, k, open0, ty);
            T inv = T(1.0) / r;
            u = jx * inv;
            v = jy * inv;
            w = (jz + T(0.5) * force_z) * inv;
        }
    } else {
        T inv = T(1.0) / r;
        u = jx * inv;
        v = jy * inv;
        w = jz * inv;
    }
    if (c.top_drive && k == c.nz - 1) {
        u = c.u_top_lb;
        v = T(0);
        w = T(0);
    }
}

// water runs also pack each cell's type and open flag into one byte for the
// stream's per-link source states (tstate; NULL otherwise)
template <class S, class T>
__global__ __launch_bounds__(NT) void k_macro(Consts<T> c, const S* f, T* rho, T* ux, T* uy, T* uz, const int* type, const uint8_t* flags,
                                              uint8_t* tstate) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    if (tstate) tstate[id] = (uint8_t)(type[id] * 2 + (int)(flags[id] & 1u));
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    T fl[19];
#pragma unroll
    for (int q = 0; q < 19; q++) fl[q] = Store<S, T>::load(f, (size_t)q * n + id, q, c.store);
    T r, u, v, w;
    macro_cell(c, k, fl, c.fs_water ? type[id] : LBM_FS_LIQUID, (flags[id] & 1u) != 0, r, u, v, w);
    rho[id] = r;
    ux[id] = u;
    uy[id] = v;
    uz[id] = w;
}

// the macro step of a water run whose rho and u already hold the macro step
// of the populations as the last stream wrote them (VOF tracking, no
// conditioning): macro_cell depends on the populations, the open flag, the
// plane and whether the cell is gas, so only the cells that changed between
// gas and wet since that stream (type_prev holds the types it saw) and the
// cells whose populations were rewritten afterwards (marked in dirty) are
// recomputed; every other cell already holds what k_macro would write
template <class S, class T>
__global__ __launch_bounds__(NT) void k_macro_fix(Consts<T> c, const S* f, T* rho, T* ux, T* uy, T* uz, const int* type,
                                                  const int* type_prev, const uint8_t* flags, uint8_t* dirty, uint8_t* tstate) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    const int ty = type[id];
    const uint8_t fl0 = flags[id];
    // the stream's source state of this cell (bits 0-2), and in bits 3-4 the
    // step ends in a row, up to 3, at which it was an open GAS cell
    const int was = tstate[id] >> 3;
    const int streak = (ty == LBM_FS_GAS && (fl0 & 1u)) ? (was < 3 ? was + 1 : 3) : 0;
    tstate[id] = (uint8_t)(ty * 2 + (int)(fl0 & 1u) + (streak << 3));
    const bool d = dirty[id] != 0;
    if ((ty == LBM_FS_GAS) == (type_prev[id] == LBM_FS_GAS) && !d) return;
    if (d) dirty[id] = 0;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    T fl[19];
#pragma unroll
    for (int q = 0; q < 19; q++) fl[q] = Store<S, T>::load(f, (size_t)q * n + id, q, c.store);
    T r, u, v, w;
    macro_cell(c, k, fl, ty, (fl0 & 1u) != 0, r, u, v, w);
    rho[id] = r;
    ux[id] = u;
    uy[id] = v;
    uz[id] = w;
}

// collision model a kernel is compiled for: KBC without top sponge and
// backscatter, regularized, or any (chosen per cell at run time)
enum { CM_ANY = 0, CM_KBC = 1, CM_REG = 2 };
static int collision_model(const lbm_config_t& cfg) {
    return cfg.collision == LBM_COLL_REGULARIZED ? CM_REG : (!cfg.enable_top_sponge && !(cfg.c_backscatter > 0.0f)) ? CM_KBC : CM_ANY;
}

// waves per SIMD the staged stream is compiled for: 4 with FP32 arithmetic
// (128 VGPRs; 129 would leave 3), no bound with FP64 (which would spill).
// (The staged regularized collision at 4, 128 VGPRs with 24 B of scratch,
// ran 116 against 100 us on the VOF workload; the KBC one needs 105.)
template <class T>
struct StagedOcc {
    static constexpr int v = sizeof(T) == 4 ? 4 : 1;
};

// compute_LES (CODEX:3011-3141) of one active cell: omega, nu_t and the
// Smagorinsky term from the velocities of its six neighbours
template <class T>
__device__ __forceinline__ T les_cell(const Consts<T>& c, const Slot<T>& s, const T* ux, const T* uy, const T* uz, const uint8_t* flags,
                                      size_t id, int i, int j, int k, T& nt, T& sm) {
    int nx = c.nx, ny = c.ny, nz = c.nz;
    int ip = wrap1(i + 1, nx), im = wrap1(i - 1, nx), jp = wrap1(j + 1, ny), jm = wrap1(j - 1, ny);
    int kp = k + 1 >= nz ? nz - 1 : k + 1, km = k - 1 < 0 ? 0 : k - 1;
    size_t pl = (size_t)nx * ny;
    size_t base = (size_t)k * pl;
    size_t xp = base + (size_t)j * nx + ip, xm = base + (size_t)j * nx + im;
    size_t yp = base + (size_t)jp * nx + i, ym = base + (size_t)jm * nx + i;
    size_t zp = (size_t)kp * pl + (size_t)j * nx + i, zm = (size_t)km * pl + (size_t)j * nx + i;
    bool near_wall = k > 0 && !(flags[id - pl] & 1u);
    T phi = T(0), uwx = T(0), uwy = T(0);
    if (near_wall) {
        phi = cell_phi(c, s, i, j, k);
        if (c.include_bed) {
            uwx = s.uw[(size_t)j * nx + i];
            uwy = s.vw[(size_t)j * nx + i];
        }
    }
    return les_omega(c, ux[xp], ux[xm], ux[yp], ux[ym], ux[zp], ux[zm], uy[xp], uy[xm], uy[yp], uy[ym], uy[zp], uy[zm], uz[xp], uz[xm],
                     uz[yp], uz[ym], uz[zp], uz[zm], near_wall, phi, ux[id], uy[id], uwx, uwy, &nt, &sm);
}

// compute_LES (CODEX:3011-3141)
template <class T>
__global__ __launch_bounds__(NT) void k_les(Consts<T> c, Slot<T> s, const T* ux, const T* uy, const T* uz, T* omega, T* nu_t, T* smag,
                      const int* type, const uint8_t* flags) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    bool active = flags[id] & 1u;
    if (c.fs_water && type[id] == LBM_FS_GAS) active = false;
    if (!active) return;
    T nt, sm;
    T om = les_cell(c, s, ux, uy, uz, flags, id, i, j, k, nt, sm);
    omega[id] = om;
    nu_t[id] = nt;
    smag[id] = sm;
}

// collide_regularized / collide_KBC (CODEX:3143-3416)
// LES: compute_LES of the cell first (as k_les: active cells only; the
// others keep their omega), then the collision with that omega
template <class S, class T, bool LES, int CM>
__global__ __launch_bounds__(NT) void k_collide(Consts<T> c, Slot<T> s, S* f, const T* rho, const T* ux, const T* uy, const T* uz,
                                                T* omega, T* nu_t, T* smag, const int* type, const uint8_t* flags, uint32_t step) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    const uint8_t fl0 = flags[id];
    bool open0 = fl0 & 1u;
    int ty = LBM_FS_LIQUID;
    if (c.fs_water) {
        ty = type[id];
        // an open GAS cell with every link open: the stream that follows
        // gives it the rest equilibrium and takes the free-surface closure
        // for the links from it, and reads its own post-collision values
        // only for the wall correction of a cell next to a wall (near, bit 1),
        // so nothing reads what its collision would store: skipped
        if (ty == LBM_FS_GAS && open0 && !(fl0 & 2u)) return;
    }
    T r = rho[id], u = ux[id], v = uy[id], w = uz[id];
    bool gas = false;
    if (c.fs_water) {
        if (ty == LBM_FS_GAS) {
            gas = true;
            r = c.rho0;
            u = v = w = T(0);
        }
    }
    T fl[19];
#pragma unroll
    for (int q = 0; q < 19; q++) fl[q] = Store<S, T>::load(f, (size_t)q * n + id, q, c.store);
    T om;
    if (LES && open0 && !gas) {
        T nt, sm;
        om = les_cell(c, s, ux, uy, uz, flags, id, i, j, k, nt, sm);
        omega[id] = om;
        nu_t[id] = nt;
        smag[id] = sm;
    } else {
        om = omega[id];
    }
    if (CM == CM_REG || (CM == CM_ANY && c.collision_reg)) {
        if (gas) {
#pragma unroll
            for (int q = 0; q < 19; q++) fl[q] = feq(r, u, v, w, q);
        } else if (c.fs_water) {
            T force_z = -r * local_g(c, k, open0, ty);
            collide_regularized<T, true>(fl, r, u, v, w, om, force_z);
        } else {
            collide_regularized<T, false>(fl, r, u, v, w, om, T(0));
        }
    } else {
        if (CM == CM_ANY && c.top_sponge && k > c.nz - 3) {
            sponge_relax(fl, r, u, v, w);
        } else {
            bool bs = CM == CM_ANY && c.c_backscatter > T(0) && open0;
            float xi = bs ? bs_uniform(c.seed, step, (uint32_t)id) : 0.0f;
            if (c.fs_water && !gas) {
                T force_z = -r * local_g(c, k, open0, ty);
                collide_kbc<T, true>(fl, r, u, v, w, om, force_z, bs, c.c_backscatter, xi);
            } else {
                collide_kbc<T, false>(fl, r, u, v, w, om, T(0), bs, c.c_backscatter, xi);
            }
        }
    }
    if constexpr (std::is_same<S, T>::value) {
#pragma unroll
        for (int q = 0; q < 19; q++) f[(size_t)q * n + id] = fl[q];
    } else {  // the store point of the fused step too (same dither)
        const uint32_t base = dither_base(step, (uint32_t)id);
#pragma unroll
        for (int q = 0; q < 19; q++) Store<S, T>::store(f, (size_t)q * n + id, q, fl[q], c.store, base);
    }
}

// free_surface_hydrostatic_density_at_cell (CODEX:821-847): one pressure
// target per node; component-balanced: the head of the still-water
// reference where gravity is balanced; hydrostatic-balanced: the surface
// elevation h - base (p' = rho g eta, may be negative); total pressure: the
// depth below the column's surface, at least 0
template <class T>
__device__ __forceinline__ T fs_density_cell(const Consts<T>& c, int k, bool open0, int type, const T* fs_h, size_t col) {
    const T dx = c.dx;
    const bool hbp = c.pf == LBM_PF_HYDROSTATIC_BALANCED || c.pf == LBM_PF_VOF_COMPONENT_BALANCED;
    T head;
    if (c.pf == LBM_PF_VOF_COMPONENT_BALANCED) {
        T zc = ((T)k + T(0.5)) * dx;
        head = gravity_fraction(c, k, open0, type) < T(0.5) ? zc - c.free_surface_base_m : T(0);
    } else if (hbp) {
        head = fs_h[col] - c.free_surface_base_m;
    } else {
        T zc = ((T)k + T(0.5)) * dx;
        head = fs_h[col] - zc;
    }
    if (head < T(0) && !hbp) head = T(0);
    return c.rho0 + c.g_lb * (head / dx) / (T(1.0) / T(3.0));
}

// free-surface gas density at a link (CODEX:821-867)
template <class T>
__device__ __forceinline__ T fs_density_link(const Consts<T>& c, int k, int q, bool open0, int type, const T* fs_h, size_t col) {
    T dx = c.dx;
    if (c.fs_water && c.pf == LBM_PF_VOF_COMPONENT_BALANCED) {
        T z_if = ((T)k + T(0.5)) * dx - T(0.5) * (T)czq(q) * dx;
        T head = T(0);
        if (gravity_fraction(c, k, open0, type) < T(0.5)) head += z_if - c.free_surface_base_m;
        return c.rho0 + c.g_lb * (head / dx) / (T(1.0) / T(3.0));
    }
    if (c.fs_water && c.pf == LBM_PF_HYDROSTATIC_BALANCED && c.fs_vof) {
        T z_if = ((T)k + T(0.5)) * dx - T(0.5) * (T)czq(q) * dx;
        T head = z_if - c.free_surface_base_m;
        return c.rho0 + c.g_lb * (head / dx) / (T(1.0) / T(3.0));
    }
    return fs_density_cell(c, k, open0, type, fs_h, col);
}

// post-collision value of population q at (si,sj,sk), with the ghost layers
// of the inflow/outflow x boundary (new): inflow = feq(rho0, u_in),
// outflow = copy of the last interior layer
template <class S, class T>
__device__ __forceinline__ T src_value(const Consts<T>& c, const S* f, size_t n, const T* inflow, int si, int sj, int sk, int q) {
    if (si < 0) return feq(c.rho0, inflow[(size_t)sk * c.ny + sj], T(0), T(0), q);
    if (si >= c.nx) si = c.nx - 1;
    size_t id = ((size_t)sk * c.ny + sj) * c.nx + si;
    return Store<S, T>::load(f, (size_t)q * n + id, q, c.store);
}

// stream + apply_x_solid_wall_bounceback + apply_open_boundary_conditions
// (CODEX:3418-3712) for one destination cell; fin = post-collision state.
// COUNT: add the link counters; MACRO: also write rho and u of the streamed
// populations (the macro step that follows stream in a water run, computed
// from the values this thread stores; each cell reads only its own velocity
// before that write)
// stream (CODEX:3418-3551), x walls, fringe, top drive and IBB
// (CODEX:3554-3712) of cell id: the post-stream populations fo, the
// cell's free-surface type and plane, and the link counts of the step
template <class S, class T>
__device__ __forceinline__ void stream_bc_cell(const Consts<T>& c, const Slot<T>& s, const S* fin, const T* ux, const T* uy, const T* uz,
                                               const uint8_t* tstate,
                                               const int* type, const T* fill, const T* u_inlet, const T* inflow, const T* fs_h,
                                               const uint8_t* flags, size_t id, T fo[19], int& ty_out, int& k_out, int& n_cut,
                                               int& n_fs, int& n_bridge) {
    size_t n = (size_t)c.nx * c.ny * c.nz;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    k_out = k;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const bool water = c.fs_water;
    const bool xper = c.x_boundary == LBM_XB_PERIODIC, xio = c.x_boundary == LBM_XB_INFLOW_OUTFLOW;
    const bool yper = c.y_boundary == LBM_YB_PERIODIC;
    const Codec& inv = c.store;
    CellGeo<T> g;
    cell_geometry(c, s, i, j, k, g);
    bool open0 = g.open_mask & 1u;
    int ty = water ? type[id] : LBM_FS_LIQUID;
    bool dest_gas = water && ty == LBM_FS_GAS && open0;
    ty_out = ty;
    T fpre[19];
    // a gas cell takes the rest equilibrium on every link; its own
    // populations are read only by the wall correction of a cell next to a wall
    if (!dest_gas || g.near) {
#pragma unroll
        for (int q = 0; q < 19; q++) fpre[q] = Store<S, T>::load(fin, (size_t)q * n + id, q, inv);
    } else {
#pragma unroll
        for (int q = 0; q < 19; q++) fpre[q] = T(0);
    }
    n_cut = 0;
    n_fs = 0;
    n_bridge = 0;
    fo[0] = dest_gas ? feq(c.rho0, T(0), T(0), T(0), 0) : fpre[0];
    size_t col = (size_t)j * nx + i;
    // source position of link q: periodic wrap, else the raw index with
    // `fluid` false outside a solid x or y wall (an inflow/outflow x
    // boundary keeps fluid and the raw index)
    auto source = [&](int q, int& si, int& sj, int& sk, bool& fluid) {
        const int sir = i - cxq(q);
        si = sir;
        fluid = true;
        if (xper)
            si = wrap1(sir, nx);
        else if (!xio && (sir < 0 || sir >= nx))
            fluid = false;
        const int sjr = j - cyq(q);
        sj = yper ? wrap1(sjr, ny) : sjr;
        if (!yper && (sjr < 0 || sjr >= ny)) fluid = false;
        sk = k - czq(q);
    };
    if (water && dest_gas) {
        // gas destination: rest equilibrium on every link
#pragma unroll
        for (int q = 1; q < 19; q++) fo[q] = feq(c.rho0, T(0), T(0), T(0), q);
    } else if (water) {
        // pass 1: the source state and population of every link, from
        // in-grid addresses (a source outside the grid reads the cell itself,
        // and the rules below ignore that value), issued before any rule so
        // their latencies overlap; the fill of an interface source is read
        // where the aperture needs it
        const bool aperture = c.fs_vof && c.link_aperture;
        // per link: 2 * type + open of the source, 3 bits each
        uint64_t st_bits = 0;
        T sv_a[19];
#pragma unroll
        for (int q = 1; q < 19; q++) {
            int si, sj, sk;
            bool fluid;
            source(q, si, sj, sk, fluid);
            const bool src_ok = fluid && sk >= 0 && sk < nz;
            const bool inx = si >= 0 && si < nx;
            const int sic = si < 0 ? 0 : (si >= nx ? nx - 1 : si);  // outflow: copy of the last layer
            const size_t sidx = src_ok ? ((size_t)sk * ny + sj) * nx + sic : id;
            const size_t tid = (src_ok && inx) ? sidx : id;
            st_bits |= (uint64_t)(tstate[tid] & 7u) << (3 * q);  // 2 type + open of the source
            sv_a[q] = Store<S, T>::load(fin, (size_t)q * n + sidx, q, inv);
        }
        // pass 2: the rules of stream (CODEX:3438-3551) for a liquid or
        // interface destination
#pragma unroll
        for (int q = 1; q < 19; q++) {
            const int qb = oppq(q);
            int si, sj, sk;
            bool fluid;
            source(q, si, sj, sk, fluid);
            const bool inz = sk >= 0 && sk < nz;
            const bool inx = si >= 0 && si < nx;
            const int stq = (int)(st_bits >> (3 * q)) & 7;
            const int sty = stq >> 1;
            const bool sopen = (stq & 1) != 0;
            const bool src_gas = (fluid && inz) ? (inx && sty == LBM_FS_GAS && sopen) : (sk >= nz);
            T out;
            if (!fluid) {
                out = fpre[qb];
                n_cut++;
            } else if (inz && !src_gas) {
                T frac = g.lf[q];
                const T sv = si < 0 ? feq(c.rho0, inflow[(size_t)sk * ny + sj], T(0), T(0), q) : sv_a[q];
                T fstr = frac * sv + (T(1.0) - frac) * fpre[qb];
                out = fstr;
                if (aperture) {
                    T la = T(1);
                    if (inx && sopen) {
                        if (sty == LBM_FS_GAS)
                            la = T(0);
                        else if (sty == LBM_FS_INTERFACE) {
                            la = fill[((size_t)sk * ny + sj) * nx + si];  // an open interface source inside the grid
                            if (la < T(0)) la = T(0);
                            if (la > T(1)) la = T(1);
                        }
                    }
                    if (la < T(1.0)) {
                        T rfs = c.rho0;
                        if (c.pf != LBM_PF_TOTAL || c.fs_boundary_hydro) rfs = fs_density_link(c, k, q, open0, ty, fs_h, col);
                        T fa = feq(rfs, ux[id], uy[id], uz[id], q);
                        T fb = feq(rfs, ux[id], uy[id], uz[id], qb);
                        T ffs = fa + fb - fpre[qb];
                        out = la * fstr + (T(1.0) - la) * ffs;
                        n_fs++;
                    }
                }
                if (frac < T(1.0)) n_cut++;
            } else if (sk < 0) {
                out = fpre[qb];
                n_cut++;
            } else {
                T rfs = c.rho0;
                if (c.pf != LBM_PF_TOTAL || c.fs_boundary_hydro) rfs = fs_density_link(c, k, q, open0, ty, fs_h, col);
                T fa = feq(rfs, ux[id], uy[id], uz[id], q);
                T fb = feq(rfs, ux[id], uy[id], uz[id], qb);
                T ffs = fa + fb - fpre[qb];
                out = ffs;
                if (c.fs_vof && c.thin_gap_bridge && sk >= 0 && sk < nz && src_gas) {
                    int bir = si - cxq(q), bi = bir;
                    int bj = yper ? pmod(sj - cyq(q), ny) : sj - cyq(q);
                    int bk = sk - czq(q);
                    bool bv = true;
                    if (xper)
                        bi = pmod(bir, nx);
                    else if (bir < 0 || bir >= nx)
                        bv = false;
                    if (!yper && (bj < 0 || bj >= ny)) bv = false;
                    if (bk < 0 || bk >= nz) bv = false;
                    if (bv) {
                        size_t bid = ((size_t)bk * ny + bj) * nx + bi;
                        if ((flags[bid] & 1u) && type[bid] != LBM_FS_GAS) {
                            T bw = fill[id];
                            if (fill[bid] < bw) bw = fill[bid];
                            bw *= c.bridge_strength;
                            if (bw > T(0)) {
                                if (bw > T(1)) bw = T(1);
                                out = (T(1.0) - bw) * ffs + bw * Store<S, T>::load(fin, (size_t)q * n + bid, q, inv);
                                n_bridge++;
                            }
                        }
                    }
                }
                n_fs++;
            }
            fo[q] = out;
        }
    } else {
#pragma unroll
        for (int q = 1; q < 19; q++) {
            const int qb = oppq(q);
            int si, sj, sk;
            bool fluid;
            source(q, si, sj, sk, fluid);
            T out;
            if (!fluid) {
                out = fpre[qb];
                n_cut++;
            } else if (sk >= 0 && sk < nz) {
                T frac = g.lf[q];
                out = frac * src_value<S, T>(c, fin, n, inflow, si, sj, sk, q) + (T(1.0) - frac) * fpre[qb];
                if (frac < T(1.0)) n_cut++;
            } else if (sk >= nz && c.top_boundary == LBM_TOP_FREE_SLIP) {
                out = src_value<S, T>(c, fin, n, inflow, si, sj, nz - 1, mirzq(q));
            } else {
                out = fpre[q];
            }
            fo[q] = out;
        }
    }
    // (A) x fringe (CODEX:3620-3644)
    if (c.x_fringe && i < 5 && open0 && k > nz / 2) {
        T sfr = T(1.0) - (T)i / T(4.0);
        T kr = (T)k / (T)(nz - 1);
        T strength = T(0.05) * (kr * kr - T(0.25));
        T sigma = strength * smoothstep01(sfr);
        T uxt = u_inlet[k] + T(0.0);
#pragma unroll
        for (int q = 0; q < 19; q++) fo[q] = (T(1.0) - sigma) * fo[q] + sigma * feq(c.rho0, uxt, T(0), T(0), q);
    }
    // (B) top drive (CODEX:3650-3662)
    if (c.top_drive && k == nz - 1 && open0) {
#pragma unroll
        for (int q = 0; q < 19; q++) fo[q] = feq(c.rho0, c.u_top_lb, T(0), T(0), q);
    }
    // (C) moving-wall IBB (CODEX:3674-3712)
    if (g.near) {
#pragma unroll
        for (int q = 1; q < 19; q++) fo[q] = ibb_link(c, s, g, i, j, q, fo[q], fpre[oppq(q)]);
    }
}

// D: the destination's type: the arithmetic type (materialised downloads)
// or the storage type (the staged step, whose 16-bit stores round with the
// dither of step ^ SALT_STREAM; the macro step then sees the stored values)
template <class S, class D, class T, bool COUNT, bool MACRO>
__global__ __launch_bounds__(NT, StagedOcc<T>::v) void k_stream_bc(Consts<T> c, Slot<T> s, const S* fin, D* fout, const T* ux, const T* uy, const T* uz,
                            const uint8_t* tstate, const int* type, const T* fill, const T* u_inlet, const T* inflow, const T* fs_h,
                            const uint8_t* flags, int* icnt, T* mrho, T* mux, T* muy, T* muz, uint32_t step, int gas_skip) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    T fo[19];
    int ty, k, n_cut, n_fs, n_bridge;
    stream_bc_cell<S, T>(c, s, fin, ux, uy, uz, tstate, type, fill, u_inlet, inflow, fs_h, flags, id, fo, ty, k, n_cut, n_fs, n_bridge);
    // gas_skip (the staged VOF step over a static geometry, storage in the
    // arithmetic type): an open GAS cell away from walls that was one at the
    // last three step ends already holds what this store would write.  Its
    // populations in fout are the rest equilibrium that the stream of two
    // steps ago stored there (the collision in between skips such a cell),
    // and rho and u the rest state the last stream wrote; a gas cell counts
    // no links.
    if (gas_skip && ty == LBM_FS_GAS && (flags[id] & 3u) == 1u && (tstate[id] >> 3) >= 3) return;
    if constexpr (std::is_same<D, T>::value) {
#pragma unroll
        for (int q = 0; q < 19; q++) fout[(size_t)q * n + id] = fo[q];
    } else {
        const uint32_t base = dither_base(step ^ SALT_STREAM, (uint32_t)id);
#pragma unroll
        for (int q = 0; q < 19; q++) fo[q] = Store<D, T>::store_rt(fout, (size_t)q * n + id, q, fo[q], c.store, base);
    }
    if (MACRO) {
        T r, u, v, w;
        macro_cell(c, k, fo, ty, (flags[id] & 1u) != 0, r, u, v, w);
        mrho[id] = r;
        mux[id] = u;
        muy[id] = v;
        muz[id] = w;
    }
    if (COUNT) {
        if (n_cut) atomicAdd(&icnt[IC_BOUNDARY_CUT], n_cut);
        if (n_fs) atomicAdd(&icnt[IC_FS_CUT], n_fs);
        if (n_bridge) atomicAdd(&icnt[IC_FS_BRIDGE], n_bridge);
    }
}

// number of cut links of the air stream (the boundary_cut_link_count of
// CODEX:3429-3549) from geometry alone, for runs on the fused path
template <class T>
__global__ __launch_bounds__(NT) void k_cut_links(Consts<T> c, Slot<T> s, int* icnt) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    int cnt = 0;
    if (id < n) {
        int i, j, k;
        cell_ijk(c, id, i, j, k);
        CellGeo<T> g;
        cell_geometry(c, s, i, j, k, g);
        for (int q = 1; q < 19; q++) {
            int sir = i - cxq(q), sjr = j - cyq(q), sk = k - czq(q);
            bool fluid = true;
            if (c.x_boundary == LBM_XB_SOLID && (sir < 0 || sir >= c.nx)) fluid = false;
            if (c.y_boundary == LBM_YB_SOLID && (sjr < 0 || sjr >= c.ny)) fluid = false;
            if (!fluid)
                cnt++;
            else if (sk >= 0 && sk < c.nz && g.lf[q] < T(1.0))
                cnt++;
        }
    }
    if (cnt) atomicAdd(&icnt[IC_BOUNDARY_CUT], cnt);
}

// ===========================================================================
// VOF pipeline (CODEX:1305-2374), deterministic gather formulation
// ===========================================================================
template <class T>
struct VofArgs {
    Consts<T> c;
    T *rho, *ux, *uy, *uz;
    T *fill, *mass, *mass2, *mdelta, *excess, *excess2;
    T *sa, *sb, *sc, *sd, *se, *smv;
    int *type, *type_prev, *excess_dir, *ext_gas, *si;
    const uint8_t* flags;
    uint8_t* mdirty;  // cells whose populations the repair rewrote (for k_macro_fix)
    int* rowdon;      // [2][nz * ny]: rows holding a donor of the redistribution round of each slot
    int* act;         // [nz * ny]: rows marked active by this step's passes (the narrow band below)
    const int* act_prev;  // the marks of the previous step's passes
    int* iso;         // [nz * ny]: rows holding an open cell with no open neighbour (set on steps with band 0)
    int band;         // 1: passes skip the rows the band shows settled; 0: every row is processed
    int* icnt;
    double* dcnt;
    const T* fs_h;  // column surface height
};

__device__ __forceinline__ bool nbr(int nx, int ny, int nz, int xb, int yb, int i, int j, int k, int di, int dj, int dk, int& ni, int& nj,
                                    int& nk) {
    int a = i + di, b = j + dj, e = k + dk;
    bool valid = true;
    if (xb == LBM_XB_PERIODIC)
        a = pmod(a, nx);
    else if (a < 0 || a >= nx)
        valid = false;
    if (yb == LBM_YB_PERIODIC)
        b = pmod(b, ny);
    else if (b < 0 || b >= ny)
        valid = false;
    if (e < 0 || e >= nz) valid = false;
    ni = a;
    nj = b;
    nk = e;
    return valid;
}
#define NBR(di, dj, dk, ni, nj, nk) nbr(a.c.nx, a.c.ny, a.c.nz, a.c.x_boundary, a.c.y_boundary, i, j, k, di, dj, dk, ni, nj, nk)
#define CID(kk, jj, ii) (((size_t)(kk) * a.c.ny + (jj)) * a.c.nx + (ii))
// index of neighbour (di, dj, dk), or of the cell itself where that neighbour
// lies beyond a wall (valid false): always in the grid, so the loads of
// every link can be issued unconditionally and combined afterwards
template <class T>
__device__ __forceinline__ size_t nbr_id(const Consts<T>& c, int i, int j, int k, int di, int dj, int dk, size_t self, bool& valid) {
    // inside the grid on every axis (the usual case): a fixed offset
    if ((unsigned)(i + di) < (unsigned)c.nx && (unsigned)(j + dj) < (unsigned)c.ny && (unsigned)(k + dk) < (unsigned)c.nz) {
        valid = true;
        return self + (ptrdiff_t)di + (ptrdiff_t)dj * c.nx + (ptrdiff_t)dk * ((ptrdiff_t)c.nx * c.ny);
    }
    int ni, nj, nk;
    valid = nbr(c.nx, c.ny, c.nz, c.x_boundary, c.y_boundary, i, j, k, di, dj, dk, ni, nj, nk);
    return valid ? ((size_t)nk * c.ny + nj) * c.nx + ni : self;
}
#define NBR_ID(di, dj, dk, valid) nbr_id(a.c, i, j, k, di, dj, dk, id, valid)

// over links q0 .. q1-1: add the wet neighbours (open, not gas) to wet and
// those of them with fill >= thr to res (loads issued unconditionally)
template <class T>
__device__ __forceinline__ void wet_resolved(const VofArgs<T>& a, int i, int j, int k, size_t id, int q0, int q1, T thr, int& wet,
                                             int& res) {
#pragma unroll
    for (int q = q0; q < q1; q++) {
        bool v;
        const size_t nid = NBR_ID(cxq(q), cyq(q), czq(q), v);
        const bool w = v & ((a.flags[nid] & 1u) != 0) & (a.type[nid] != LBM_FS_GAS);
        wet += w;
        res += w & (a.fill[nid] >= thr);
    }
}
#define VOF_PROLOGUE                                                      \
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;            \
    size_t n = (size_t)a.c.nx * a.c.ny * a.c.nz;                          \
    if (id >= n) return;                                                  \
    int i, j, k;                                                          \
    cell_ijk(a.c, id, i, j, k);

// Narrow band.  A cell is settled when it is LIQUID or GAS (with no mass),
// holds what the classification of its own mass gives (mass, fill, type),
// with no excess and no excess direction, and no pass of this step has
// changed it otherwise.  Every pass that leaves a cell unsettled, changes its type, or
// changes its mass, fill or type outside a classification marks the cell's
// row (k, j) in act; the preclassification, which classifies every cell,
// marks from scratch each step.  For a settled cell the passes below are
// no-ops, except for the stores of values it already holds:
//   the classification rewrites the same mass, fill and type, excess 0;
//   the redistribution has no donor there (a donor has excess);
//   the interface layer converts a LIQUID cell next to a GAS cell, and the
//   orphan and detached tests act on a cell whose wet neighbours are few
//   or unresolved: each needs an unsettled cell or a LIQUID-GAS contact
//   among the 18 neighbours, i.e. an active row within one row.
// LIQUID-GAS contacts: the previous step ended with the interface layer
// enforced, which leaves none, and the types of settled cells did not
// change since, so every contact has a marked cell.  What this does not
// cover runs with band = 0 (every row processed): the first step after the
// state is set from outside (its marks are not known) and a moving geometry
// (cells open and close).  A LIQUID cell with no open
// neighbour has no resolved one either: the orphan test always processes
// the rows holding such a cell (iso, found on the steps with band = 0).
// The scratch si of a cell is nonzero only where a pass marked it (donors,
// orphans, movers), and the preclassification clears it everywhere.
template <class T>
__device__ __forceinline__ void band_mark(const VofArgs<T>& a, int j, int k) {
    a.act[(size_t)k * a.c.ny + j] = 1;
}
// rows (k, j) and the 8 around it (y wrapped as the neighbours are; a row
// beyond a z end or a solid y wall is replaced by (k, j) itself): the nine
// loads are issued together, with no branch between them
template <class T>
__device__ __forceinline__ bool rows_near(const VofArgs<T>& a, const int* rows, int j, int k) {
    const int ny = a.c.ny, nz = a.c.nz;
    const int km = k > 0 ? k - 1 : k, kp = k + 1 < nz ? k + 1 : k;
    int jm, jp;
    if (a.c.y_boundary == LBM_YB_PERIODIC) {
        jm = wrap1(j - 1, ny);
        jp = wrap1(j + 1, ny);
    } else {
        jm = j > 0 ? j - 1 : j;
        jp = j + 1 < ny ? j + 1 : j;
    }
    const int* r0 = rows + (size_t)km * ny;
    const int* r1 = rows + (size_t)k * ny;
    const int* r2 = rows + (size_t)kp * ny;
    return (r0[jm] | r0[j] | r0[jp] | r1[jm] | r1[j] | r1[jp] | r2[jm] | r2[j] | r2[jp]) != 0;
}
// row (k, j) may hold a cell that a pass changes: itself active, or (near)
// one of its 8 neighbouring rows active
template <class T>
__device__ __forceinline__ bool band_row(const VofArgs<T>& a, int j, int k, bool near) {
    if (!a.band) return true;
    if (!near) return a.act[(size_t)k * a.c.ny + j] != 0;
    return rows_near(a, a.act, j, k);
}

// compute_vof_mass_exchange (CODEX:1305-1350)
template <class S, class T>
__global__ __launch_bounds__(NT) void k_vof_exchange(VofArgs<T> a, const S* f) {
    VOF_PROLOGUE
    if (id == 0) {  // the per-step counters of the pipeline (each is only added to after this pass)
        a.icnt[IC_VOF_EXCESS] = 0;
        a.icnt[IC_VOF_ORPHAN] = 0;
        a.icnt[IC_VOF_ADVECT] = 0;
        a.icnt[IC_VOF_COMPRESS] = 0;
        a.icnt[IC_VOF_COLLAPSE_CELLS] = 0;
        a.icnt[IC_FS_REFILL] = 0;
        a.icnt[IC_FS_REPAIR] = 0;
        for (int t = 0; t < 2; t++) a.dcnt[DC_ADVECT_MASS + t] = 0.0;  // advected and compressed mass
        for (int t = 0; t < 5; t++) a.dcnt[DC_CAND_MASS + t] = 0.0;    // the collapse sums
    }
    if (i == 0) {
        a.act[(size_t)k * a.c.ny + j] = 0;  // this step's marks start here
        if (!a.band) a.iso[(size_t)k * a.c.ny + j] = 0;
    }
    // every INTERFACE cell was marked where it became one (the band's marks
    // of the previous step cover its end state)
    if (a.band && a.act_prev[(size_t)k * a.c.ny + j] == 0) return;
    // the delta is read only for these cells (the preclassification, which
    // also sets every cell's excess and its direction)
    if (!((a.flags[id] & 1u) && a.type[id] == LBM_FS_INTERFACE)) return;
    T dm = T(0);
    const T fill_c = a.fill[id];
    // links in two halves of 9: the loads of a half are all issued before
    // its sums (in q order)
#pragma unroll
    for (int h = 0; h < 2; h++) {
        bool wet[9];
        int ty[9];
        T fn[9], fa[9], fb[9];
#pragma unroll
        for (int e = 0; e < 9; e++) {
            const int q = 1 + 9 * h + e;
            bool v;
            const size_t nid = NBR_ID(cxq(q), cyq(q), czq(q), v);
            wet[e] = v & ((a.flags[nid] & 1u) != 0);
            ty[e] = a.type[nid];
            fn[e] = a.fill[nid];
            fa[e] = Store<S, T>::load(f, (size_t)oppq(q) * n + id, oppq(q), a.c.store);
            fb[e] = Store<S, T>::load(f, (size_t)q * n + nid, q, a.c.store);
        }
#pragma unroll
        for (int e = 0; e < 9; e++) {
            if (wet[e] & (ty[e] != LBM_FS_GAS)) {
                T lw = T(1.0);
                if (ty[e] == LBM_FS_INTERFACE) lw = T(0.5) * (fill_c + fn[e]);
                if (lw > T(0)) dm += lw * (fa[e] - fb[e]);
            }
        }
    }
    a.mdelta[id] = dm;
}

// apply_vof_mass_update_and_preclassify (CODEX:1353-1414)
template <class T>
__global__ __launch_bounds__(NT) void k_vof_preclassify(VofArgs<T> a, T empty_thr) {
    VOF_PROLOGUE
    if (id == 0) {  // the next redistribution round starts without donors
        a.icnt[IC_REDIST_DONORS] = 0;
        a.icnt[IC_REDIST_DONORS + 1] = 0;
    }
    clear_donor_rows(a, 0, i, j, k);
    clear_donor_rows(a, 1, i, j, k);
    int old = a.type[id];
    // a row the last step left unmarked holds settled cells only, LIQUID or
    // GAS: type_prev, the excess and its direction, si, a GAS cell's mass and
    // fill and a LIQUID cell's fill already hold what this pass writes (the
    // type did not change since the last preclassification, and no excess
    // was left); a LIQUID cell's mass follows its density
    if (a.band && a.act_prev[(size_t)k * a.c.ny + j] == 0) {
        if (old == LBM_FS_LIQUID) {
            T rr = a.rho[id];
            if (rr < T(1.0e-8)) rr = a.c.rho0;
            a.mass[id] = rr;
        }
        return;
    }
    a.type_prev[id] = old;
    a.excess[id] = T(0);
    a.excess_dir[id] = 0;
    a.si[id] = 0;  // nonzero only at cells the last step marked (its flood-fill flags are cleared by the collapse)
    if (!(a.flags[id] & 1u)) {
        a.mass[id] = T(0);
        a.fill[id] = T(0);
        a.type[id] = LBM_FS_GAS;
        if (old != LBM_FS_GAS) band_mark(a, j, k);
        return;
    }
    if (!a.band) {  // (the geometry is the one of every later step with band 1)
        bool any = false;
#pragma unroll
        for (int q = 1; q < 19; q++) {
            bool v;
            const size_t nid = NBR_ID(cxq(q), cyq(q), czq(q), v);
            any |= v & ((a.flags[nid] & 1u) != 0);
        }
        if (!any) a.iso[(size_t)k * a.c.ny + j] = 1;
    }
    T rr = a.rho[id];
    if (rr < T(1.0e-8)) rr = a.c.rho0;
    T m = a.mass[id];
    if (old == LBM_FS_INTERFACE)
        m += a.mdelta[id];
    else if (old == LBM_FS_LIQUID)
        m = rr;
    else
        m = T(0);
    T ex = T(0);
    int ed = 0;
    if (m < T(0)) {
        ex = m;
        ed = -1;
        m = T(0);
    } else if (m > rr) {
        ex = m - rr;
        ed = 1;
        m = rr;
    }
    T fl = m / rr;
    int ct = LBM_FS_INTERFACE;
    if (fl > T(1.0e-4) && fl < empty_thr) {
        ex = m;
        ed = -1;
        m = T(0);
        fl = T(0);
        ct = LBM_FS_GAS;
    } else if (fl <= T(1.0e-4)) {
        fl = T(0);
        ct = LBM_FS_GAS;
    } else if (fl >= T(1.0) - T(1.0e-4)) {
        fl = T(1.0);
        ct = LBM_FS_LIQUID;
    }
    a.mass[id] = m;
    a.excess[id] = ex;
    a.excess_dir[id] = ed;
    if (fabs(ex) > T(1.0e-7)) atomicAdd(&a.icnt[IC_VOF_EXCESS], 1);
    a.fill[id] = fl;
    a.type[id] = ct;
    if (ct == LBM_FS_INTERFACE || ct != old || ex != T(0)) band_mark(a, j, k);
}

template <class T>
__device__ __forceinline__ T vof_room(const VofArgs<T>& a, size_t nid, T ex) {
    if (ex > T(0)) {
        T rr = a.rho[nid];
        if (rr < T(1.0e-8)) rr = a.c.rho0;
        return rr - a.mass[nid];
    }
    return a.mass[nid];
}

// redistribute_vof_mass_excess (CODEX:1417-1574) phase A: donor sums
template <class T>
__global__ __launch_bounds__(NT) void k_vof_redist_a(VofArgs<T> a, int slot) {
    VOF_PROLOGUE
    // a donor has excess (marked); si is already zero in a settled row
    if (!band_row(a, j, k, false)) return;
    a.si[id] = 0;
    T ex = a.excess[id];
    if (!(fabs(ex) > T(1.0e-12))) return;
    int ed = a.excess_dir[id];
    if (ed == 0) ed = ex < T(0) ? -1 : 1;
    int nx = a.c.nx, ny = a.c.ny, nz = a.c.nz;
    int ip = i + 1, im = i - 1, jp = j + 1, jm = j - 1;
    if (a.c.x_boundary == LBM_XB_PERIODIC) {
        ip = wrap1(i + 1, nx);
        im = wrap1(i - 1, nx);
    } else {
        if (ip >= nx) ip = nx - 1;
        if (im < 0) im = 0;
    }
    if (a.c.y_boundary == LBM_YB_PERIODIC) {
        jp = wrap1(j + 1, ny);
        jm = wrap1(j - 1, ny);
    } else {
        if (jp >= ny) jp = ny - 1;
        if (jm < 0) jm = 0;
    }
    int kp = k + 1 >= nz ? nz - 1 : k + 1, km = k - 1 < 0 ? 0 : k - 1;
    T nxv = -(a.fill[CID(k, j, ip)] - a.fill[CID(k, j, im)]);
    T nyv = -(a.fill[CID(k, jp, i)] - a.fill[CID(k, jm, i)]);
    T nzv = -(a.fill[CID(kp, j, i)] - a.fill[CID(km, j, i)]);
    T nm = sqrt(nxv * nxv + nyv * nyv + nzv * nzv);
    if (nm > T(1.0e-8)) {
        T invn = T(1.0) / nm;
        nxv *= invn;
        nyv *= invn;
        nzv *= invn;
    }
    // every neighbour's flags, type, density and mass loaded before the sums
    // (the cell itself stands in for a neighbour beyond a wall)
    bool nv[19];
    uint8_t nfl[19];
    int nty[19];
    T nrho[19], nmass[19];
#pragma unroll
    for (int q = 1; q < 19; q++) {
        const size_t nid = NBR_ID(cxq(q), cyq(q), czq(q), nv[q]);
        nfl[q] = a.flags[nid];
        nty[q] = a.type[nid];
        nrho[q] = a.rho[nid];
        nmass[q] = a.mass[nid];
    }
    T ws = T(0), cs = T(0);
#pragma unroll
    for (int q = 1; q < 19; q++) {
        if (!nv[q] || !(nfl[q] & 1u)) continue;
        T nd = (T)cxq(q) * nxv + (T)cyq(q) * nyv + (T)czq(q) * nzv;
        bool ok = ed > 0 ? nty[q] != LBM_FS_LIQUID : nty[q] != LBM_FS_GAS;
        if (!ok) continue;
        T room = nmass[q];  // vof_room
        if (ex > T(0)) room = (nrho[q] < T(1.0e-8) ? a.c.rho0 : nrho[q]) - nmass[q];
        if (room > T(1.0e-12)) {
            cs += room;
            T wgt = ed > 0 ? nd : -nd;
            if (wgt > T(0)) ws += wgt * room;
        }
    }
    a.si[id] = ed;
    a.sa[id] = nxv;
    a.sb[id] = nyv;
    a.sc[id] = nzv;
    a.sd[id] = ws;
    a.se[id] = cs;
    a.icnt[IC_REDIST_DONORS + slot] = 1;  // every donor stores the same value
    a.rowdon[(size_t)slot * a.c.nz * a.c.ny + (size_t)k * a.c.ny + j] = 1;
}

// phase B: the mass of cell id after the round: its own mass, its own
// retained excess, then the shares of its donors in q order
template <class T>
__device__ __forceinline__ T redist_gather(const VofArgs<T>& a, int i, int j, int k, size_t id) {
    T acc = a.mass[id];
    if (a.si[id] != 0 && !(a.se[id] > T(1.0e-12))) acc += a.excess[id];
    if (a.flags[id] & 1u) {
        int ty = a.type[id];
#pragma unroll
        for (int q = 1; q < 19; q++) {
            bool v;
            const size_t sid = NBR_ID(-cxq(q), -cyq(q), -czq(q), v);
            const int ed = a.si[sid];
            const T se = a.se[sid];
            if (!v | (ed == 0) | !(se > T(1.0e-12))) continue;
            T ex = a.excess[sid];
            bool ok = ed > 0 ? ty != LBM_FS_LIQUID : ty != LBM_FS_GAS;
            if (!ok) continue;
            T room = vof_room(a, id, ex);
            if (!(room > T(1.0e-12))) continue;
            T nd = (T)cxq(q) * a.sa[sid] + (T)cyq(q) * a.sb[sid] + (T)czq(q) * a.sc[sid];
            T wgt = ed > 0 ? nd : -nd;
            T share = ex * room / a.se[sid];
            if (a.sd[sid] > T(1.0e-12) && wgt > T(0))
                share = ex * wgt * room / a.sd[sid];
            else if (a.sd[sid] > T(1.0e-12))
                share = T(0);
            acc += share;
        }
    }
    return acc;
}

// finalize_vof_classification_from_mass (CODEX:1577-1624) of cell id whose
// mass is m: mass, fill, type, excess (into ex_out) and its direction
template <class T>
__device__ __forceinline__ void finalize_cell(const VofArgs<T>& a, size_t id, int j, int k, T m, T empty_thr, T* ex_out) {
    if (!(a.flags[id] & 1u)) {
        ex_out[id] = T(0);
        a.excess_dir[id] = 0;
        a.mass[id] = T(0);
        a.fill[id] = T(0);
        a.type[id] = LBM_FS_GAS;
        return;
    }
    T rr = a.rho[id];
    if (rr < T(1.0e-8)) rr = a.c.rho0;
    T ex = T(0);
    int ed = 0, cnt = 0;
    if (m < T(0)) {
        ex = m;
        ed = -1;
        m = T(0);
        cnt++;
    }
    if (m > rr) {
        ex = m - rr;
        ed = 1;
        m = rr;
        cnt++;
    }
    T fl = m / rr;
    int ct = LBM_FS_INTERFACE;
    if (fl > T(1.0e-4) && fl < empty_thr) {
        ex = m;
        ed = -1;
        m = T(0);
        fl = T(0);
        ct = LBM_FS_GAS;
        cnt++;
    } else if (fl <= T(1.0e-4)) {
        fl = T(0);
        ct = LBM_FS_GAS;
    } else if (fl >= T(1.0) - T(1.0e-4)) {
        fl = T(1.0);
        ct = LBM_FS_LIQUID;
    }
    const int old = a.type[id];
    a.mass[id] = m;
    ex_out[id] = ex;
    a.excess_dir[id] = ed;
    a.fill[id] = fl;
    a.type[id] = ct;
    if (cnt) atomicAdd(&a.icnt[IC_VOF_EXCESS], cnt);
    // (a GAS cell that kept a sliver of mass is not settled: the next
    // preclassification zeroes it)
    if (ct == LBM_FS_INTERFACE || ct != old || cnt || (ct == LBM_FS_GAS && m != T(0))) band_mark(a, j, k);
}

// one redistribution round's phase B followed by the classification
// (redistribute_vof_mass_excess + finalize_vof_classification_from_mass).
// A cell's gather reads its own mass, type and excess and its donors'
// excess and weights (phase A's arrays), so the mass is updated in place and
// the excess is written to the other buffer (excess2), which the host then
// swaps in.  The donor flag of a round is slot `slot` of two; this kernel
// clears the other one, which the next round's phase A sets.
// Rows of donors: a cell with no donor in its own row or the 8 rows around
// it (the rows of its 18 neighbours, y wrapped as the neighbours are) has
// no share to gather and keeps its mass.
template <class T>
__device__ __forceinline__ void clear_donor_rows(const VofArgs<T>& a, int slot, int i, int j, int k) {
    if (i == 0) a.rowdon[(size_t)slot * a.c.nz * a.c.ny + (size_t)k * a.c.ny + j] = 0;
}

template <class T>
__global__ __launch_bounds__(NT) void k_vof_redist_fin(VofArgs<T> a, T empty_thr, int slot) {
    VOF_PROLOGUE
    if (id == 0) a.icnt[IC_REDIST_DONORS + (slot ^ 1)] = 0;
    clear_donor_rows(a, slot ^ 1, i, j, k);
    // no donor this round, or none within reach: the mass is unchanged
    const int* rows = a.rowdon + (size_t)slot * a.c.nz * a.c.ny;
    const bool gather = a.icnt[IC_REDIST_DONORS + slot] != 0 && rows_near(a, rows, j, k);
    // a settled cell with no share: the classification gives excess 0 and
    // rewrites what the cell holds
    if (!gather && !band_row(a, j, k, false)) {
        a.excess2[id] = T(0);
        return;
    }
    const T m = gather ? redist_gather(a, i, j, k, id) : a.mass[id];
    finalize_cell(a, id, j, k, m, empty_thr, a.excess2);
}

// enforce_vof_interface_layer (CODEX:1627-1656): only LIQUID cells change
// and only GAS neighbours are read, so the in-place update has no race
template <class T>
__global__ __launch_bounds__(NT) void k_vof_enforce(VofArgs<T> a) {
    VOF_PROLOGUE
    if (!band_row(a, j, k, true)) return;  // a LIQUID-GAS contact has a marked cell
    if (!((a.flags[id] & 1u) && a.type[id] == LBM_FS_LIQUID)) return;
    bool touches = false;
#pragma unroll
    for (int q = 1; q < 19; q++) {
        bool v;
        const size_t nid = NBR_ID(cxq(q), cyq(q), czq(q), v);
        touches |= v & ((a.flags[nid] & 1u) != 0) & (a.type[nid] == LBM_FS_GAS);
    }
    if (touches) {
        a.type[id] = LBM_FS_INTERFACE;
        a.fill[id] = T(1.0);
        a.mass[id] = a.rho[id];
        band_mark(a, j, k);
    }
}

// remove_orphan_vof_interface_cells (CODEX:1659-1776) phase A: decide
template <class T>
__global__ __launch_bounds__(NT) void k_vof_orphan_a(VofArgs<T> a, T empty_thr, T thr, int rmax, int max_weak) {
    VOF_PROLOGUE
    // a settled LIQUID cell with no unsettled or GAS neighbour has a
    // resolved one, unless it has no open neighbour at all (iso)
    if (!band_row(a, j, k, true) && a.iso[(size_t)k * a.c.ny + j] == 0) return;
    a.si[id] = 0;
    if (!((a.flags[id] & 1u) && a.type[id] != LBM_FS_GAS && a.fill[id] > empty_thr)) return;
    int wet = 0, res = 0;
    wet_resolved(a, i, j, k, id, 1, 7, thr, wet, res);  // the 6 face neighbours
    if (res != 0) return;                                // a resolved neighbour: not an orphan
    wet_resolved(a, i, j, k, id, 7, 19, thr, wet, res);  // the 12 edge neighbours
    if (!(res == 0 && (a.fill[id] < thr || wet <= max_weak))) return;
    T m = a.mass[id];
    T cw = T(0), tw = T(0);
    for (int q = 1; q < 19; q++)
        for (int d = 2; d <= rmax; d++) {
            int ni, nj, nk;
            int di = d * cxq(q), dj = d * cyq(q), dk = d * czq(q);
            if (!NBR(di, dj, dk, ni, nj, nk)) continue;
            size_t nid = CID(nk, nj, ni);
            if ((a.flags[nid] & 1u) && a.type[nid] != LBM_FS_GAS && a.fill[nid] >= thr) {
                T rr = a.rho[nid];
                if (rr < T(1.0e-8)) rr = a.c.rho0;
                T room = rr - a.mass[nid];
                T d2 = (T)(di * di + dj * dj + dk * dk);
                tw += a.fill[nid] / d2;
                if (room > T(1.0e-12)) cw += room / d2;
            }
        }
    if (m > T(1.0e-12) && (cw > T(1.0e-12) || tw > T(1.0e-12))) {
        a.si[id] = 1;
        a.sb[id] = cw;
        a.sc[id] = tw;
        atomicAdd(&a.icnt[IC_VOF_ORPHAN], 1);
        band_mark(a, j, k);
    }
}

// phase B: receivers gather the scattered orphan mass into mass2
template <class T>
__global__ __launch_bounds__(NT) void k_vof_orphan_b(VofArgs<T> a, T thr, int rmax) {
    VOF_PROLOGUE
    if (a.icnt[IC_VOF_ORPHAN] == 0) return;  // phase A found no orphan (the common case): nothing moves
    bool orphan = a.si[id] != 0;
    T acc = orphan ? T(0) : a.mass[id];
    if ((a.flags[id] & 1u) && a.type[id] != LBM_FS_GAS && a.fill[id] >= thr) {
        T rr = a.rho[id];
        if (rr < T(1.0e-8)) rr = a.c.rho0;
        T room = rr - a.mass[id];
        for (int q = 1; q < 19; q++)
            for (int d = 2; d <= rmax; d++) {
                int si_, sj, sk;
                int di = d * cxq(q), dj = d * cyq(q), dk = d * czq(q);
                if (!NBR(-di, -dj, -dk, si_, sj, sk)) continue;
                size_t sid = CID(sk, sj, si_);
                if (!a.si[sid]) continue;
                T d2 = (T)(di * di + dj * dj + dk * dk);
                T wgt = T(0), den = a.sb[sid];
                if (a.sb[sid] > T(1.0e-12)) {
                    if (room > T(1.0e-12)) wgt = room / d2;
                } else {
                    wgt = a.fill[id] / d2;
                    den = a.sc[sid];
                }
                if (den > T(1.0e-12) && wgt > T(0)) acc += a.mass[sid] * wgt / den;
            }
    }
    a.mass2[id] = acc;
}

// phase C: the gathered masses replace the masses and the orphans become gas
template <class T>
__global__ __launch_bounds__(NT) void k_vof_orphan_c(VofArgs<T> a) {
    VOF_PROLOGUE
    if (a.icnt[IC_VOF_ORPHAN] == 0) return;
    const T m = a.mass2[id];
    const bool moved = m != a.mass[id];
    a.mass[id] = m;
    if (a.si[id]) {
        a.fill[id] = T(0);
        a.type[id] = LBM_FS_GAS;
    }
    if (moved || a.si[id]) band_mark(a, j, k);
}

// vof_unresolved_detached_cell (CODEX:1779-1810)
template <class T>
__device__ bool vof_detached(const VofArgs<T>& a, int i, int j, int k, T orphan_thr, int max_wet, int max_res) {
    size_t id = CID(k, j, i);
    if (!((a.flags[id] & 1u) && a.type[id] != LBM_FS_GAS && a.fill[id] > T(1.0e-4))) return false;
    int wet = 0, res = 0;
    wet_resolved(a, i, j, k, id, 1, 7, orphan_thr, wet, res);  // the 6 face neighbours
    if (wet > max_wet || res > max_res) return false;           // the counts only grow: decided
    wet_resolved(a, i, j, k, id, 7, 19, orphan_thr, wet, res);  // the 12 edge neighbours
    return wet <= max_wet && res <= max_res;
}

__device__ __constant__ int FD6[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

// advect_detached_vof_cells / compress_detached_vof_residuals phase A
// (CODEX:1813-1993): mass each detached cell moves through each face
template <class T>
__global__ __launch_bounds__(NT) void k_vof_detached_a(VofArgs<T> a, int mode, T orphan_thr, int max_wet, int max_res, T resid_thr) {
    VOF_PROLOGUE
    // a mover has a GAS face neighbour: a settled mover would be a contact
    if (!band_row(a, j, k, true)) return;
    T mv[6] = {T(0), T(0), T(0), T(0), T(0), T(0)};
    if (vof_detached(a, i, j, k, orphan_thr, max_wet, max_res)) {
        T m = a.mass[id];
        T u = a.ux[id], v = a.uy[id], w = a.uz[id];
        T sg[6] = {u, -u, v, -v, w, -w};
        T r[6];
        for (int d = 0; d < 6; d++) {
            r[d] = T(0);
            if (sg[d] > T(0)) {
                int ni, nj, nk;
                if (NBR(FD6[d][0], FD6[d][1], FD6[d][2], ni, nj, nk)) {
                    size_t nid = CID(nk, nj, ni);
                    if ((a.flags[nid] & 1u) && a.type[nid] == LBM_FS_GAS) r[d] = sg[d];
                }
            }
        }
        if (mode == 0) {
            if (m > T(1.0e-12)) {
                T rs = r[0] + r[1] + r[2] + r[3] + r[4] + r[5];
                if (rs > T(1.0e-12)) {
                    T scale = T(1.0);
                    if (rs > T(0.95)) scale = T(0.95) / rs;
                    for (int d = 0; d < 6; d++)
                        if (r[d] > T(0)) {
                            T x = m * r[d] * scale;
                            mv[d] = x > T(1.0e-12) ? x : T(0);
                        }
                }
            }
        } else {
            T rr = a.rho[id];
            if (rr < T(1.0e-8)) rr = a.c.rho0;
            T fl = m / rr;
            if (fl > T(1.0e-4) && fl < resid_thr) {
                int best = -1;
                T br = T(0);
                for (int d = 0; d < 6; d++)
                    if (r[d] > br) {
                        br = r[d];
                        best = d;
                    }
                if (br > T(1.0e-12)) {
                    T x = m * T(1.0) * T(1.0);
                    mv[best] = x > T(1.0e-12) ? x : T(0);
                }
            }
        }
    }
    T moved = T(0);
    for (int d = 0; d < 6; d++) moved += mv[d];
    // every nonzero share exceeds 1e-12, so a cell that moved no more than
    // that moved nothing: si marks the movers, and only they store shares
    const bool mover = moved > T(1.0e-12);
    a.si[id] = mover ? 1 : 0;
    if (mover) {
        band_mark(a, j, k);
        for (int d = 0; d < 6; d++) a.smv[6 * id + d] = mv[d];
        atomicAdd(&a.icnt[mode == 0 ? IC_VOF_ADVECT : IC_VOF_COMPRESS], 1);
        atomicAdd(&a.dcnt[mode == 0 ? DC_ADVECT_MASS : DC_COMPRESS_MASS], (double)moved);
    }
}

// finalize_vof_classification_from_mass (CODEX:1577-1624) after a detached
// pass (mode 0 advection, 1 compaction; -1: none ran): its phase B, the
// moves out of the cell then into it in fixed direction order, then the
// classification of the cell's own fields, in place (phase B reads only the
// movers' shares, which this kernel does not write)
template <class T>
__global__ __launch_bounds__(NT) void k_vof_detached_fin(VofArgs<T> a, T empty_thr, int mode) {
    VOF_PROLOGUE
    if (id == 0) {  // the next redistribution round starts without donors
        a.icnt[IC_REDIST_DONORS] = 0;
        a.icnt[IC_REDIST_DONORS + 1] = 0;
    }
    clear_donor_rows(a, 0, i, j, k);
    clear_donor_rows(a, 1, i, j, k);
    // phase A moved something (a cell counted): movers are marked and their
    // receivers are face neighbours
    const bool moved = mode >= 0 && a.icnt[mode == 0 ? IC_VOF_ADVECT : IC_VOF_COMPRESS] != 0;
    // a settled cell with no mover near: its excess (in place) is already zero
    if (!band_row(a, j, k, moved)) return;
    T m = a.mass[id];
    if (moved) {
        // the shares of a cell that moved nothing are all zero (not stored)
        T out = T(0);
        if (a.si[id])
            for (int d = 0; d < 6; d++) out -= a.smv[6 * id + d];
        T in = T(0);
        for (int d = 0; d < 6; d++) {
            int si_, sj, sk;
            if (!NBR(-FD6[d][0], -FD6[d][1], -FD6[d][2], si_, sj, sk)) continue;
            const size_t sid = CID(sk, sj, si_);
            if (a.si[sid]) in += a.smv[6 * sid + d];
        }
        T dm = out + in;
        if (fabs(dm) > T(1.0e-12)) {
            m = m + dm;
            band_mark(a, j, k);
        }
    }
    finalize_cell(a, id, j, k, m, empty_thr, a.excess);
}

// vof_gas_fraction_at (CODEX:1996-2009)
template <class T>
__device__ __forceinline__ T gas_fraction(const VofArgs<T>& a, size_t id) {
    T g = T(0);
    if (a.flags[id] & 1u) {
        if (a.type[id] == LBM_FS_GAS)
            g = T(1.0);
        else if (a.fill[id] < T(1.0)) {
            g = T(1.0) - a.fill[id];
            if (g < T(0)) g = T(0);
            if (g > T(1)) g = T(1);
        }
    }
    return g;
}

// initialize_vof_exterior_gas_marks (CODEX:2021-2030)
template <class T>
__global__ __launch_bounds__(NT) void k_flood_seed(VofArgs<T> a) {
    VOF_PROLOGUE
    a.ext_gas[id] = (gas_fraction(a, id) > T(1.0e-6) && k == a.c.nz - 1) ? 1 : 0;
}

// propagate_vof_exterior_gas_marks (CODEX:2033-2055), run to the fixed
// point: every sweep after a sweep that marked nothing returns at once
template <class T>
__global__ __launch_bounds__(NT) void k_flood_sweep(VofArgs<T> a, int* changed_prev, int* changed_this) {
    if (*changed_prev == 0) return;
    // cells in decreasing index order: the seeds are the top plane, and the
    // marks of the blocks that finished earlier in the sweep (higher planes)
    // carry on down within the same sweep (the fixed point does not depend
    // on the order)
    const size_t n = (size_t)a.c.nx * a.c.ny * a.c.nz;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t id = n - 1 - t;
    int i, j, k;
    cell_ijk(a.c, id, i, j, k);
    if (a.ext_gas[id] != 0 || !(gas_fraction(a, id) > T(1.0e-6))) return;
    bool conn = false;
#pragma unroll
    for (int q = 1; q < 7; q++) {
        int ni, nj, nk;
        if (!NBR(cxq(q), cyq(q), czq(q), ni, nj, nk)) continue;
        if (a.ext_gas[CID(nk, nj, ni)] == 1) conn = true;
    }
    if (conn) {
        a.ext_gas[id] = 1;
        *changed_this = 1;
    }
}

// summarize_vof_trapped_gas_collapse sums (CODEX:2058-2075) in FP64
// candidate mass, candidate volume and exterior capacity of the trapped
// gas (CODEX:2059-2105), as per-chunk FP64 partials (three per chunk,
// interleaved); k_collapse_decide sums them in a fixed order
template <class T>
__global__ __launch_bounds__(RED_THREADS) void k_collapse_sum(VofArgs<T> a, double* part) {
    const size_t n = (size_t)a.c.nx * a.c.ny * a.c.nz;
    const size_t base = (size_t)blockIdx.x * RED_CHUNK + threadIdx.x;
    double cm = 0.0, cv = 0.0, cap = 0.0;
    for (int r = 0; r < RED_PER_THREAD; r++) {
        const size_t id = base + (size_t)r * RED_THREADS;
        if (id >= n) break;
        T g = gas_fraction(a, id);
        if (g > T(1.0e-6)) {
            if (a.ext_gas[id] == 0) {
                T rr = a.rho[id];
                if (rr < T(1.0e-8)) rr = a.c.rho0;
                cm += (double)(g * rr);
                cv += (double)(g * a.c.dx * a.c.dx * a.c.dx);
            } else if (a.type[id] != LBM_FS_GAS) {
                T m = a.mass[id];
                if (m > T(0)) cap += (double)m;
            }
        }
    }
    cm = block_tree<false>(cm);
    cv = block_tree<false>(cv);
    cap = block_tree<false>(cap);
    if (threadIdx.x == 0) {
        part[3 * (size_t)blockIdx.x] = cm;
        part[3 * (size_t)blockIdx.x + 1] = cv;
        part[3 * (size_t)blockIdx.x + 2] = cap;
    }
}
__global__ __launch_bounds__(RED_THREADS) void k_collapse_decide(const double* part, int count, double* dcnt, double max_volume) {
    double s3[3];
    for (int e = 0; e < 3; e++) {
        double v = 0.0;
        for (int i = (int)threadIdx.x; i < count; i += RED_THREADS) v += part[3 * (size_t)i + e];
        s3[e] = block_tree<false>(v);
    }
    if (threadIdx.x != 0) return;
    const double cm = s3[0], cv = s3[1], cap = s3[2];
    dcnt[DC_CAND_MASS] = cm;
    dcnt[DC_CAND_VOLUME] = cv;
    dcnt[DC_EXT_CAPACITY] = cap;
    dcnt[DC_APPLIED_VOLUME] = 0.0;
    dcnt[DC_REMOVAL_SCALE] = 0.0;
    if (cv > 0.0 && cv <= max_volume && cap >= cm && cap > 1.0e-12) {
        dcnt[DC_APPLIED_VOLUME] = cv;
        dcnt[DC_REMOVAL_SCALE] = cm / cap;
    }
}
template <class T>
__global__ __launch_bounds__(NT) void k_collapse_apply(VofArgs<T> a, int nflags) {
    VOF_PROLOGUE
    if (id < (size_t)nflags) a.si[id] = 0;  // the flood fill's sweep flags (kept in si)
    T scale = (T)a.dcnt[DC_REMOVAL_SCALE];
    if (!(scale > T(0))) return;
    T g = gas_fraction(a, id);
    if (!(g > T(1.0e-6))) return;
    T rr = a.rho[id];
    if (rr < T(1.0e-8)) rr = a.c.rho0;
    if (a.ext_gas[id] == 0) {
        a.mass[id] = rr;
        a.fill[id] = T(1.0);
        a.type[id] = LBM_FS_LIQUID;
        atomicAdd(&a.icnt[IC_VOF_COLLAPSE_CELLS], 1);
        band_mark(a, j, k);
    } else if (a.type[id] != LBM_FS_GAS) {
        T m = a.mass[id] * (T(1.0) - scale);
        if (m < T(0)) m = T(0);
        T fl = m / rr;
        int ct = LBM_FS_INTERFACE;
        if (fl <= T(1.0e-4)) {
            fl = T(0);
            m = T(0);
            ct = LBM_FS_GAS;
        } else if (fl >= T(1.0) - T(1.0e-4)) {
            fl = T(1.0);
            m = rr;
            ct = LBM_FS_LIQUID;
        }
        a.mass[id] = m;
        a.fill[id] = fl;
        a.type[id] = ct;
        band_mark(a, j, k);
    }
}

// update_free_surface_height_from_vof (CODEX:2146-2161): the height of a
// column is the bed height plus the sum over k, in k order, of
// fill * centre fraction * dx over its open cells.  A block takes HC
// columns: its threads stage the terms of HK planes at a time in LDS (every
// load of the chunk in flight at once, and the centre fractions computed in
// parallel; a closed cell's term is +0, which leaves the sum unchanged: it
// starts at +0 and every term is >= 0), then one thread per column adds
// them in k order, 16 LDS reads ahead of the adds.
constexpr int HC = 32;    // columns per block
constexpr int HK = 192;   // planes per staged chunk
constexpr int HT = 256;   // threads per block
template <class T>
__global__ __launch_bounds__(HT) void k_vof_height(Consts<T> c, Slot<T> s, const T* fill, const uint8_t* flags, T* fs_h) {
    __shared__ T add_s[HK][HC];
    const size_t plane = (size_t)c.nx * c.ny;
    const int lc = (int)threadIdx.x % HC, lr = (int)threadIdx.x / HC;
    const size_t col = (size_t)blockIdx.x * HC + lc;
    const bool colok = col < plane;
    const int i = colok ? (int)(col % (size_t)c.nx) : 0, j = colok ? (int)(col / (size_t)c.nx) : 0;
    // the bed alone sets the centre fraction: cell_center_fraction with the
    // column's bed height read once
    const bool bed_only = !c.include_cube && !c.phi_fraction;
    const T eta_c = (bed_only && colok) ? s.eta9[4 * plane + col] : T(0);
    T depth = T(0);
    for (int k0 = 0; k0 < c.nz; k0 += HK) {
        const int kn = c.nz - k0 < HK ? c.nz - k0 : HK;
        // every load of the chunk issued before any term (planes beyond the
        // grid and columns beyond the plane read a clamped in-grid cell;
        // their terms are not read back)
        constexpr int E = HK / (HT / HC);
        T fl[E];
        bool op[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int kk = lr + e * (HT / HC);
            const int k = k0 + kk < c.nz ? k0 + kk : c.nz - 1;
            const size_t id = (size_t)k * plane + (colok ? col : 0);
            op[e] = flags[id] & 1u;
            fl[e] = fill[id];
        }
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int kk = lr + e * (HT / HC);
            const int k = k0 + kk < c.nz ? k0 + kk : c.nz - 1;
            T frac;
            if (bed_only) {
                const T zc = ((T)k + T(0.5)) * c.dx;
                const T gap = c.include_bed ? zc - eta_c : T(1.0e6);
                frac = T(0.5) + div_dx(c, gap);
                if (frac < T(0)) frac = T(0);
                if (frac > T(1)) frac = T(1);
            } else {
                bool by_cube;
                frac = cell_center_fraction(c, s, i, j, k, by_cube);
            }
            add_s[kk][lc] = op[e] ? fl[e] * frac * c.dx : T(0);
        }
        __syncthreads();
        if (lr == 0) {
            int kk = 0;
            for (; kk + 16 <= kn; kk += 16) {
                T v[16];
#pragma unroll
                for (int e = 0; e < 16; e++) v[e] = add_s[kk + e][lc];
#pragma unroll
                for (int e = 0; e < 16; e++) depth = depth + v[e];
            }
            for (; kk < kn; kk++) depth = depth + add_s[kk][lc];
        }
        __syncthreads();
    }
    if (lr == 0 && colok) fs_h[col] = s.h[col] + depth;
}

// ---- height-function free-surface tracking (CODEX:1209-1232, 2376-2429) ----

// classify_free_surface_cells: fill of the column height in cell k, types by
// the 1e-4 thresholds, closed cells gas; the old type moves to type_prev
template <class T>
__global__ __launch_bounds__(NT) void k_fs_classify(Consts<T> c, const T* fs_h, const uint8_t* flags, T* fill, int* type, int* type_prev) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    const int old = type[id];
    T z_bottom = (T)k * c.dx;
    T fl = (fs_h[(size_t)j * c.nx + i] - z_bottom) / c.dx;
    if (fl < T(0.0)) fl = T(0.0);
    if (fl > T(1.0)) fl = T(1.0);
    int t = LBM_FS_INTERFACE;
    if (fl <= T(1e-4))
        t = LBM_FS_GAS;
    else if (fl >= T(1.0) - T(1e-4))
        t = LBM_FS_LIQUID;
    if (!(flags[id] & 1u)) {
        fl = T(0.0);
        t = LBM_FS_GAS;
    }
    fill[id] = fl;
    type_prev[id] = old;
    type[id] = t;
}

// compute_free_surface_column_flux: q = sum over the open, non-gas cells of
// the column of fill u vel_scale dx, bottom to top
template <class T>
__global__ __launch_bounds__(NT) void k_fs_flux(Consts<T> c, const T* fill, const int* type, const uint8_t* flags, const T* ux,
                                                const T* uy, T* qx, T* qy) {
    size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t plane = (size_t)c.nx * c.ny;
    if (col >= plane) return;
    T sx = T(0), sy = T(0);
    for (int k = 0; k < c.nz; k++) {
        size_t id = (size_t)k * plane + col;
        if ((flags[id] & 1u) && type[id] != LBM_FS_GAS) {
            T fl = fill[id];
            sx += fl * ux[id] * c.vel_scale * c.dx;
            sy += fl * uy[id] * c.vel_scale * c.dx;
        }
    }
    qx[col] = sx;
    qy[col] = sy;
}

// divergence of a column flux along one axis: central differences on a
// periodic axis, half-sum face fluxes with no flux through the walls of a
// solid one
template <class T>
__device__ __forceinline__ T flux_div(const T* q, size_t stride, int a, int n, bool solid, T dx) {
    if (solid) {
        T fl = T(0), fr = T(0);
        if (a > 0) fl = T(0.5) * (q[(size_t)(a - 1) * stride] + q[(size_t)a * stride]);
        if (a < n - 1) fr = T(0.5) * (q[(size_t)a * stride] + q[(size_t)(a + 1) * stride]);
        return (fr - fl) / dx;
    }
    int ap = a + 1 == n ? 0 : a + 1, am = a == 0 ? n - 1 : a - 1;
    return (q[(size_t)ap * stride] - q[(size_t)am * stride]) * (T(0.5) / dx);
}

// solitary_surface_height_at_x (CODEX:1166-1182): the KdV wave
// A sech^2(kappa (x - x0 - dir c t)), kappa = sqrt(3 A / (4 h0^3)), offset
// folded into the periodic domain; exp is det_exp
template <class T>
__device__ __forceinline__ T solitary_height(const Consts<T>& c, T xp, T t) {
    const T lx = (T)c.nx * c.dx;
    const T amp = c.solitary[0], h0 = c.solitary[1];
    const T kappa = sqrt(T(3.0) * amp / (T(4.0) * h0 * h0 * h0));
    const T center = c.solitary[2] + c.solitary[3] * c.solitary[4] * t;
    T xrel = xp - center;
    if (c.x_boundary == LBM_XB_PERIODIC) {
        if (xrel > T(0.5) * lx) xrel -= lx;
        if (xrel < T(-0.5) * lx) xrel += lx;
    }
    const T arg = kappa * xrel;
    const T ep = (T)det_exp((double)arg), em = (T)det_exp((double)(-arg));
    const T sech = T(2.0) / (ep + em);
    return c.free_surface_base_m + amp * sech * sech;
}

// advect_free_surface_height: h += dt (-(dq_x/dx + dq_y/dy)) between half a
// cell above the bed and half a cell below the lid (height_kinematic); the
// solitary wave at time t (prescribed_solitary); height_static keeps h.
// Solid y walls (new) take the x-wall form.
template <class T>
__global__ __launch_bounds__(NT) void k_fs_advect(Consts<T> c, const T* qx, const T* qy, const T* bed, T* fs_h, T t) {
    size_t col = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t plane = (size_t)c.nx * c.ny;
    if (col >= plane) return;
    const int i = (int)(col % (size_t)c.nx), j = (int)(col / (size_t)c.nx);
    if (c.fs_mode == LBM_FSM_PRESCRIBED_SOLITARY) {
        fs_h[col] = solitary_height(c, ((T)i + T(0.5)) * c.dx, t);
        return;
    }
    if (c.fs_mode != LBM_FSM_HEIGHT_KINEMATIC) return;
    const T dx = c.dx;
    const T h_old = fs_h[col];
    T dq_dx = flux_div(qx + (size_t)j * c.nx, 1, i, c.nx, c.x_boundary == LBM_XB_SOLID, dx);
    T dq_dy = flux_div(qy + i, (size_t)c.nx, j, c.ny, c.y_boundary == LBM_YB_SOLID, dx);
    T dh_dt = -(dq_dx + dq_dy);
    T h_new = h_old + c.dt_phys * dh_dt;
    T min_h = bed[col] + T(0.5) * dx;
    T max_h = (T)c.nz * dx - T(0.5) * dx;
    if (h_new < min_h) h_new = min_h;
    if (h_new > max_h) h_new = max_h;
    fs_h[col] = h_new;
}

// refill_new_free_surface_cells (CODEX:2163-2316)
// 16-bit storage: the refilled cell's populations are stored rounded with
// the dither of step ^ SALT_REFILL
template <class S, class T>
__global__ __launch_bounds__(NT) void k_vof_refill(VofArgs<T> a, S* f, uint32_t step) {
    VOF_PROLOGUE
    if (!band_row(a, j, k, false)) return;  // the cell's type changed this step: marked
    if (!(a.type_prev[id] == LBM_FS_GAS && a.type[id] != LBM_FS_GAS && (a.flags[id] & 1u))) return;
    const Consts<T>& c = a.c;
    T ws = 0, fws = 0, ra = 0, xa = 0, ya = 0, za = 0, rf = 0, xf = 0, yf = 0, zf = 0;
    T fna[19], fnf[19];
#pragma unroll
    for (int q = 0; q < 19; q++) fna[q] = fnf[q] = T(0);
    for (int q = 1; q < 19; q++) {
        int ni, nj, nk;
        if (!NBR(cxq(q), cyq(q), czq(q), ni, nj, nk)) continue;
        size_t nid = CID(nk, nj, ni);
        if (!((a.flags[nid] & 1u) && a.type_prev[nid] != LBM_FS_GAS && a.type[nid] != LBM_FS_GAS)) continue;
        T wn = wq<T>(q), fl = a.fill[nid];
        T rn = a.rho[nid], uxn = a.ux[nid], uyn = a.uy[nid], uzn = a.uz[nid];
        fws += wn;
        rf += wn * rn;
        xf += wn * uxn;
        yf += wn * uyn;
        zf += wn * uzn;
        T lw = wn * fl;
        ws += lw;
        ra += lw * rn;
        xa += lw * uxn;
        ya += lw * uyn;
        za += lw * uzn;
        if (c.refill_mode == LBM_REFILL_NONEQ)
            for (int qq = 0; qq < 19; qq++) {
                T fneq = Store<S, T>::load(f, (size_t)qq * n + nid, qq, c.store) - feq(rn, uxn, uyn, uzn, qq);
                fna[qq] += lw * fneq;
                fnf[qq] += wn * fneq;
            }
    }
    // VOF: density from the liquid neighbours; height tracking: the
    // hydrostatic density of the column (CODEX:2225-2246)
    const bool vof = c.fs_vof;
    // (the gravity fraction of the reference's field is the one of the step's
    // start, when this cell was still gas: type_prev)
    T rr = vof ? c.rho0 : fs_density_cell(c, k, (a.flags[id] & 1u) != 0, a.type_prev[id], a.fs_h, (size_t)j * c.nx + i);
    T urx = c.u_current_lb, ury = T(0), urz = T(0);
    if (ws > T(0)) {
        T inv = T(1.0) / ws;
        if (vof) rr = ra * inv;
        urx = xa * inv;
        ury = ya * inv;
        urz = za * inv;
    } else if (fws > T(0)) {
        T inv = T(1.0) / fws;
        if (vof) rr = rf * inv;
        urx = xf * inv;
        ury = yf * inv;
        urz = zf * inv;
    } else if (vof) {
        rr = c.rho0;
    }
    T fv[19];
    T rp = T(0);
    for (int q = 0; q < 19; q++) {
        T fr = feq(rr, urx, ury, urz, q);
        if (c.refill_mode == LBM_REFILL_NONEQ) {
            if (ws > T(0))
                fr += fna[q] / ws;
            else if (fws > T(0))
                fr += fnf[q] / fws;
        } else if (c.refill_mode == LBM_REFILL_DIRECTIONAL) {
            int si_, sj, sk;
            if (NBR(-cxq(q), -cyq(q), -czq(q), si_, sj, sk)) {
                size_t sid = CID(sk, sj, si_);
                if ((a.flags[sid] & 1u) && a.type_prev[sid] != LBM_FS_GAS && a.type[sid] != LBM_FS_GAS)
                    fr = Store<S, T>::load(f, (size_t)q * n + sid, q, c.store);
            }
        }
        fv[q] = fr;
        rp += fr;
    }
    const uint32_t base = dither_base(step ^ SALT_REFILL, (uint32_t)id);
    auto put = [&](int q, T v) {
        if constexpr (std::is_same<S, T>::value)
            f[(size_t)q * n + id] = v;
        else
            Store<S, T>::store(f, (size_t)q * n + id, q, v, c.store, base);
    };
    if (c.refill_mode == LBM_REFILL_DIRECTIONAL) {
        T scale = T(1.0);
        if (rp > T(1.0e-8)) scale = rr / rp;
        T r2 = 0, mx = 0, my = 0, mz = 0;
        for (int q = 0; q < 19; q++) {
            T fvv = fv[q] * scale;
            put(q, fvv);
            r2 += fvv;
            mx += fvv * (T)cxq(q);
            my += fvv * (T)cyq(q);
            mz += fvv * (T)czq(q);
        }
        a.rho[id] = r2;
        if (r2 > T(1.0e-8)) {
            a.ux[id] = mx / r2;
            a.uy[id] = my / r2;
            a.uz[id] = mz / r2;
        } else {
            a.ux[id] = urx;
            a.uy[id] = ury;
            a.uz[id] = urz;
        }
    } else {
        a.rho[id] = rr;
        a.ux[id] = urx;
        a.uy[id] = ury;
        a.uz[id] = urz;
        for (int q = 0; q < 19; q++) put(q, fv[q]);
    }
    atomicAdd(&a.icnt[IC_FS_REFILL], 1);
}

// repair_vof_coalescence_links (CODEX:2319-2373): donors (refilled cells,
// type_prev gas) and repaired cells (type_prev wet) are disjoint
// the copies move stored values unchanged (the shift of a population
// depends on q only); the moments are taken from the values as loaded
template <class S, class T>
__global__ __launch_bounds__(NT) void k_vof_repair(VofArgs<T> a, S* f) {
    VOF_PROLOGUE
    // a fresh source is a cell the refill just filled (the refill's own
    // condition), so a step that refilled nothing repairs nothing
    if (a.icnt[IC_FS_REFILL] == 0) return;
    if (!band_row(a, j, k, true)) return;  // next to a refilled (marked) cell
    if (!((a.flags[id] & 1u) && a.type[id] != LBM_FS_GAS && a.type_prev[id] != LBM_FS_GAS)) return;
    bool rep = false;
    int cnt = 0;
    // the sources that just filled (loads only, then the copies)
    uint32_t fresh = 0;
    size_t src[19];
#pragma unroll
    for (int q = 1; q < 19; q++) {
        bool v;
        src[q] = NBR_ID(-cxq(q), -cyq(q), -czq(q), v);
        const bool nf = v & ((a.flags[src[q]] & 1u) != 0) & (a.type_prev[src[q]] == LBM_FS_GAS) & (a.type[src[q]] != LBM_FS_GAS);
        fresh |= (uint32_t)nf << q;
    }
    for (int q = 1; q < 19; q++) {
        if ((fresh >> q) & 1u) {
            f[(size_t)q * n + id] = f[(size_t)q * n + src[q]];
            rep = true;
            cnt++;
        }
    }
    if (rep) {
        T r = 0, mx = 0, my = 0, mz = 0;
        for (int q = 0; q < 19; q++) {
            T fq = Store<S, T>::load(f, (size_t)q * n + id, q, a.c.store);
            r += fq;
            mx += fq * (T)cxq(q);
            my += fq * (T)cyq(q);
            mz += fq * (T)czq(q);
        }
        a.rho[id] = r;
        if (r > T(1.0e-8)) {
            a.ux[id] = mx / r;
            a.uy[id] = my / r;
            a.uz[id] = mz / r;
        }
        a.mdirty[id] = 1;
        atomicAdd(&a.icnt[IC_FS_REPAIR], cnt);
    }
}

// initialize_vof_mass_from_fill (CODEX:1289-1302)
template <class T>
__global__ __launch_bounds__(NT) void k_vof_init_mass(VofArgs<T> a) {
    VOF_PROLOGUE
    if ((a.flags[id] & 1u) && a.type[id] != LBM_FS_GAS) {
        a.mass[id] = a.fill[id] * a.rho[id];
    } else {
        a.mass[id] = T(0);
        a.fill[id] = T(0);
        a.type[id] = LBM_FS_GAS;
    }
    a.type_prev[id] = a.type[id];
    a.mdelta[id] = T(0);
    a.excess[id] = T(0);
    a.excess_dir[id] = 0;
}

// ===========================================================================
// fused step (air mode): moments pre-pass + fused stream-LES-collide pass
// ===========================================================================
//
// The state between steps is the post-collision population set f* of step
// n (one ping-pong buffer pair).  A step is two launches over the grid, one
// thread per cell:
//   K1 (k_moments): pull-streams f* of step n with the boundary rules of
//      step n, conditions the closed links for step n+1 and writes the
//      velocity of step n+1 (ux, uy, uz) - 76 B read + 12 B written per
//      cell (FP32);
//   K2 (k_collide): repeats the identical pull + conditioning, evaluates the
//      LES viscosity from the 6 neighbouring velocities written by K1,
//      collides and writes f* of step n+1 - 76 B + 12 B read, 76 B written.
// The LES central differences (CODEX:3040-3141) need the velocity of the
// neighbours of step n+1, which exist only after their own pull; K1 makes
// them available to K2 for 12 B/cell of extra traffic instead of a halo
// recomputation inside a tile.
//
// Planes up to the wall band (the highest plane a moving bed can cut) run
// the BAND instantiation, which evaluates the link geometry link by link
// from the 9 column bed samples; the planes above it run the lean one
// (plain pull with the domain-boundary rules only).
constexpr int FB_THREADS = 256;  // block: bx x (256 / bx) cells of one plane, bx = 64 .. 256


template <class S, class T>
struct FusedArgs {
    Consts<T> c;
    Slot<T> s0;  // time level of the pending stream (step n)
    Slot<T> s1;  // time level of the new step (n+1)
    const S* fin;
    S* fout;
    T *rho, *ux, *uy, *uz, *omega;
    const T *u_inlet, *inflow;
    uint8_t* open;  // band planes: 1 if the cell has an open link at step n+1 (K1 writes, K2 reads)
    int mode;       // 0: fin holds post-stream populations (no stream)
    int write_macro;
    int plane0;  // first plane of this launch
    int band_planes;  // k_collide_all: planes 0 .. band_planes - 1 take the band rules
    int strips;       // tiles of a plane dealt to the XCDs in strips of rows (tile_of)
    uint32_t step;
};


// c_q . u over the non-zero lattice components, in x, y, z order; a zero
// component contributes an exact zero to the reference's sum, so the value
// equals (cx ux + cy uy) + cz uz
template <class T>
__device__ __forceinline__ T cdotu(int q, T ux, T uy, T uz) {
    const int a = cxq(q), b = cyq(q), e = czq(q);
    T s = T(0);
    bool first = true;
    if (a) {
        s = a > 0 ? ux : -ux;
        first = false;
    }
    if (b) {
        T v = b > 0 ? uy : -uy;
        s = first ? v : s + v;
        first = false;
    }
    if (e) {
        T v = e > 0 ? uz : -uz;
        s = first ? v : s + v;
    }
    return s;
}

// feq with the velocity terms hoisted: u15 = 1.5 |u|^2 (CODEX:641-643)
template <class T>
__device__ __forceinline__ T feq_h(T rho, T ux, T uy, T uz, T u15, int q) {
    T cu = T(3.0) * cdotu<T>(q, ux, uy, uz);
    return rho * wq<T>(q) * (T(1.0) + cu + T(0.5) * cu * cu - u15);
}
// feq_h for all 19 q, bit for bit: the opposite link has c.u negated
// exactly, so 3 c.u, (3 c.u)^2 / 2 and rho w are evaluated once per pair
template <class T>
__device__ __forceinline__ void feq_all(T rho, T ux, T uy, T uz, T u15, T fe[19]) {
    constexpr int base[9] = {1, 3, 5, 7, 8, 11, 12, 15, 16};  // one link of each opposite pair
    fe[0] = feq_h<T>(rho, ux, uy, uz, u15, 0);
#pragma unroll
    for (int p = 0; p < 9; p++) {
        const int q = base[p], qb = oppq(q);
        const T cu = T(3.0) * cdotu<T>(q, ux, uy, uz);
        const T h = T(0.5) * cu * cu;
        const T rw = rho * wq<T>(q);
        fe[q] = rw * (T(1.0) + cu + h - u15);
        fe[qb] = rw * (T(1.0) - cu + h - u15);
    }
}

// Q_q : Pi over the non-zero entries of Q (zero entries add exact zeros)
template <class T>
__device__ __forceinline__ T q_dot_pi(int q, T Pxx, T Pyy, T Pzz, T Pxy, T Pxz, T Pyz) {
    const T cs2 = T(1.0) / T(3.0);
    const int a = cxq(q), b = cyq(q), e = czq(q);
    T d = ((T)(a * a) - cs2) * Pxx + ((T)(b * b) - cs2) * Pyy + ((T)(e * e) - cs2) * Pzz;
    T o;
    if (a * b)
        o = (T)(a * b) * Pxy;
    else if (a * e)
        o = (T)(a * e) * Pxz;
    else if (b * e)
        o = (T)(b * e) * Pyz;
    else
        return d;
    return d + T(2.0) * o;
}

// collide_regularized, air mode (CODEX:3324-3416), in place
template <class T>
__device__ __forceinline__ void fused_regularized(T f[19], T rho, T ux, T uy, T uz, T om) {
    const T cs2 = T(1.0) / T(3.0);
    const T inv2 = T(1.0) / (T(2.0) * cs2 * cs2);
    const T eps = T(1e-6);
    T u2 = ux * ux + uy * uy + uz * uz;
    T u15 = T(1.5) * u2;
    T fe[19];
    T Pxx = T(0), Pyy = T(0), Pzz = T(0), Pxy = T(0), Pxz = T(0), Pyz = T(0);
    feq_all<T>(rho, ux, uy, uz, u15, fe);
#pragma unroll
    for (int q = 0; q < 19; q++) {
        T d = f[q] - fe[q];
        const int a = cxq(q), b = cyq(q), e = czq(q);
        if (a) Pxx += d;
        if (b) Pyy += d;
        if (e) Pzz += d;
        if (a * b) Pxy += a * b > 0 ? d : -d;
        if (a * e) Pxz += a * e > 0 ? d : -d;
        if (b * e) Pyz += b * e > 0 ? d : -d;
    }
    T alpha = T(1.0);
    const T one_m = T(1.0) - om;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        T freg = wq<T>(q) * inv2 * q_dot_pi<T>(q, Pxx, Pyy, Pzz, Pxy, Pxz, Pyz);
        T fs = fe[q] + one_m * freg;
        f[q] = fs;
        if (fs < eps) {
            T den = fs - fe[q];
            if (den < T(0)) {
                T cand = (eps - fe[q]) / den;
                if (cand < alpha) alpha = cand;
            }
        }
    }
    if (alpha < T(0)) alpha = T(0);
    if (alpha > T(1)) alpha = T(1);
    if (alpha != T(1.0)) {
#pragma unroll
        for (int q = 0; q < 19; q++) f[q] = fe[q] + alpha * (f[q] - fe[q]);
    } else {
        // alpha == 1: fe + 1 * (fs - fe) rounds back to fs unless the
        // difference itself rounds; evaluate it for bit-identity
#pragma unroll
        for (int q = 0; q < 19; q++) f[q] = fe[q] + (f[q] - fe[q]);
    }
}

// collide_KBC without backscatter, air mode (CODEX:3210-3243), in place
template <class T>
__device__ __forceinline__ void fused_kbc(T f[19], T rho, T ux, T uy, T uz, T om) {
    const T eps = T(1e-6);
    T u2 = ux * ux + uy * uy + uz * uz;
    T u15 = T(1.5) * u2;
    T fe[19];
    T alpha = T(1.0);
    feq_all<T>(rho, ux, uy, uz, u15, fe);
#pragma unroll
    for (int q = 0; q < 19; q++) {
        T fs = f[q] - om * (f[q] - fe[q]);
        f[q] = fs;
        if (fs < eps) {
            T den = fs - fe[q];
            if (den < T(0)) {
                T cand = (eps - fe[q]) / den;
                if (cand < alpha) alpha = cand;
            }
        }
    }
    if (alpha < T(0)) alpha = T(0);
    if (alpha > T(1)) alpha = T(1);
    if (alpha != T(1.0)) {
#pragma unroll
        for (int q = 0; q < 19; q++) f[q] = fe[q] + alpha * (f[q] - fe[q]);
    } else {
        // 1 * x == x: the same values without the product
#pragma unroll
        for (int q = 0; q < 19; q++) f[q] = fe[q] + (f[q] - fe[q]);
    }
}

// link q of a cell in column (i,j) leaves the domain through a solid x or y
// wall (CODEX:2655-2672 for x; y walls are new)
template <class T>
__host__ __device__ __forceinline__ bool link_on_wall(const Consts<T>& c, int i, int j, int q) {
    return (c.x_boundary == LBM_XB_SOLID && ((i == 0 && cxq(q) > 0) || (i == c.nx - 1 && cxq(q) < 0))) ||
           (c.y_boundary == LBM_YB_SOLID && ((j == 0 && cyq(q) > 0) || (j == c.ny - 1 && cyq(q) < 0)));
}

// open fraction and open flag of link q of cell (i,j,k) from the column bed
// samples (CODEX:2573-2680), bed obstacle + x/y wall overrides
template <class T>
__device__ __forceinline__ T link_geo_e(const Consts<T>& c, T eta_q, int i, int j, int k, int q, int& lo, bool& wall) {
    const T dx = c.dx;
    T zc = ((T)k + T(0.5)) * dx;
    T zq = zc + T(0.5) * dx * (T)czq(q);
    T gap = zq - eta_q;
    T ggap = c.include_bed ? gap : T(1.0e6);
    T frac = T(0.5) + div_dx(c, ggap);
    if (frac < T(0)) frac = T(0);
    if (frac > T(1)) frac = T(1);
    lo = ggap < T(0) ? 0 : 1;
    wall = q > 0 && link_on_wall(c, i, j, q);
    if (wall) {
        lo = 0;
        frac = T(0);
    }
    return frac;
}

template <class T>
__device__ __forceinline__ T link_geo(const Consts<T>& c, const T* eta9, size_t col, size_t plane2, int i, int j, int k, int q, int& lo,
                                      bool& wall) {
    return link_geo_e(c, eta9[(size_t)c9q(q) * plane2 + col], i, j, k, q, lo, wall);
}

// open mask (bit q, q = 1..18) of a cell from its column's 9 endpoint heights
template <class T>
__device__ __forceinline__ uint32_t link_mask_e(const Consts<T>& c, const T e9[9], int i, int j, int k) {
    uint32_t m = 0;
#pragma unroll
    for (int q = 1; q < 19; q++) {
        int lo;
        bool w;
        link_geo_e(c, e9[c9q(q)], i, j, k, q, lo, w);
        m |= (uint32_t)lo << q;
    }
    return m;
}

// open mask (bit q, q = 1..18) of a cell
template <class T>
__device__ __forceinline__ uint32_t link_mask(const Consts<T>& c, const T* eta9, size_t col, size_t plane2, int i, int j, int k) {
    T e9[9];
#pragma unroll
    for (int e = 0; e < 9; e++) e9[e] = eta9[(size_t)e * plane2 + col];
    return link_mask_e(c, e9, i, j, k);
}

// wall type of link q (bed unless an x/y wall override)
template <class T>
__device__ __forceinline__ int link_wt(const Consts<T>& c, bool wall, int i, int j, int q) {
    if (!wall) return c.include_bed ? LBM_WALL_BED : LBM_WALL_NONE;
    if (c.y_boundary == LBM_YB_SOLID && j == 0 && cyq(q) > 0) return LBM_WALL_YMIN;
    if (c.y_boundary == LBM_YB_SOLID && j == c.ny - 1 && cyq(q) < 0) return LBM_WALL_YMAX;
    return (i == 0 && cxq(q) > 0) ? LBM_WALL_XMIN : LBM_WALL_XMAX;
}

// x fringe (CODEX:3620-3644) and top drive (CODEX:3650-3662) applied to the
// pulled populations of an open cell
template <class S, class T>
__device__ __forceinline__ void stream_edges(const FusedArgs<S, T>& a, int i, int k, bool open0, T f[19]) {
    const Consts<T>& c = a.c;
    const int nz = c.nz;
    if (c.x_fringe && i < 5 && open0 && k > nz / 2) {
        T sfr = T(1.0) - (T)i / T(4.0);
        T kr = (T)k / (T)(nz - 1);
        T strength = T(0.05) * (kr * kr - T(0.25));
        T sigma = strength * smoothstep01(sfr);
        T uxt = a.u_inlet[k] + T(0.0);
#pragma unroll
        for (int q = 0; q < 19; q++) f[q] = (T(1.0) - sigma) * f[q] + sigma * feq(c.rho0, uxt, T(0), T(0), q);
    }
    if (c.top_drive && k == nz - 1 && open0) {
#pragma unroll
        for (int q = 0; q < 19; q++) f[q] = feq(c.rho0, c.u_top_lb, T(0), T(0), q);
    }
}

// Pull of step n (+ boundary rules) and conditioning of step n+1 for a cell
// of the wall band; returns the populations entering the moments of step
// n+1 and whether lattice_open[0] holds at step n+1.  The rules follow
// stream (CODEX:3418-3551), the x walls / fringe / top drive / IBB of
// CODEX:3554-3712 and update_bed_populations_and_reinit (CODEX:2894-2940).
// A wave lies in one row (j) of one plane (k), so y and z decisions are
// uniform over it; per-lane x decisions (bed, x edges) are selects, and all
// loads of a cell are issued before the arithmetic that consumes them.
template <class S, class T>
__device__ __forceinline__ void band_gather(const FusedArgs<S, T>& a, int i, int j_lane, int k, T f[19], bool& open0_out) {
    const Consts<T>& c = a.c;
    const int j = __builtin_amdgcn_readfirstlane(j_lane);
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const uint32_t n = (uint32_t)nx * ny * nz;
    const uint32_t plane = (uint32_t)nx * ny;
    const uint32_t id = (uint32_t)k * plane + (uint32_t)j * nx + i;
    const Codec& cod = c.store;
    const uint32_t col = (uint32_t)j * nx + i;
    const S* fin = a.fin;
    if (k <= a.s1.ksolid[col]) {
        // closed at step n+1: the conditioning replaces every population by
        // feq(rho0, wall velocity) with weight 1 and the streamed value with
        // weight 0 (equal whenever the streamed values are finite)
        const T* fs = a.s1.fsol + col;
#pragma unroll
        for (int q = 0; q < 19; q++) f[q] = fs[(size_t)q * plane];
        open0_out = false;
        return;
    }
#define LD(q, idx) Store<S, T>::load(fin + (size_t)(q) * n, (uint32_t)(idx), (q), cod)
    const bool slow0 = a.mode != 0 && k <= a.s0.kband[col];
    const bool slow1 = k <= a.s1.kband[col];
    if (a.mode == 0) {
#pragma unroll
        for (int q = 0; q < 19; q++) f[q] = LD(q, id);
    } else {
        const bool xper = c.x_boundary == LBM_XB_PERIODIC, xio = c.x_boundary == LBM_XB_INFLOW_OUTFLOW;
        const bool yper = c.y_boundary == LBM_YB_PERIODIC;
        const bool free_slip = c.top_boundary == LBM_TOP_FREE_SLIP;
        // source columns / rows, clamped into the grid at non-periodic edges
        // (those links take the edge rules below)
        const int im = i == 0 ? (xper ? nx - 1 : 0) : i - 1;
        const int ip = i == nx - 1 ? (xper ? 0 : nx - 1) : i + 1;
        const int jm = j == 0 ? (yper ? ny - 1 : 0) : j - 1;
        const int jp = j == ny - 1 ? (yper ? 0 : ny - 1) : j + 1;
        // x edges: solid walls bounce back, the inflow column takes the
        // equilibrium ghost at the inflow velocity
        const bool xlo = !xper && i == 0, xhi = !xper && !xio && i == nx - 1;
        f[0] = LD(0, id);
        T loc[19];
        const bool any_slow = __builtin_amdgcn_ballot_w64(slow0) != 0;
        if (any_slow || xlo || xhi) {
#pragma unroll
            for (int q = 1; q < 19; q++) loc[q] = LD(q, id);
        }
        T e0[9];
        uint32_t mask0 = 0x7fffeu;
        if (any_slow) {
#pragma unroll
            for (int e = 0; e < 9; e++) e0[e] = a.s0.eta9[(size_t)e * plane + col];
        }
#pragma unroll
        for (int q = 1; q < 19; q++) {
            const int cxi = cxq(q), cyi = cyq(q), czi = czq(q);
            const int si = cxi > 0 ? im : (cxi < 0 ? ip : i);
            const int sj = cyi > 0 ? jm : (cyi < 0 ? jp : j);
            const bool yout = !yper && ((cyi < 0 && j == ny - 1) || (cyi > 0 && j == 0));
            const int sk = k - czi;
            if (yout)
                f[q] = LD(oppq(q), id);
            else if (sk >= 0 && sk < nz)
                f[q] = LD(q, (uint32_t)sk * plane + (uint32_t)sj * nx + si);
            else if (sk >= nz && free_slip)
                f[q] = LD(mirzq(q), (uint32_t)(nz - 1) * plane + (uint32_t)sj * nx + si);
            else
                f[q] = LD(q, id);
        }
        if (xlo || xhi) {
#pragma unroll
            for (int q = 1; q < 19; q++) {
                const int cxi = cxq(q), cyi = cyq(q), czi = czq(q);
                if (!((xlo && cxi > 0) || (xhi && cxi < 0))) continue;
                const bool yout = !yper && ((cyi < 0 && j == ny - 1) || (cyi > 0 && j == 0));
                const int sj = cyi > 0 ? jm : (cyi < 0 ? jp : j);
                const int sk = k - czi;
                if (!xio || yout)
                    f[q] = loc[oppq(q)];
                else if (sk >= 0 && sk < nz)
                    f[q] = feq(c.rho0, a.inflow[(size_t)sk * ny + sj], T(0), T(0), q);
                else if (sk >= nz && free_slip)
                    f[q] = feq(c.rho0, a.inflow[(size_t)(nz - 1) * ny + sj], T(0), T(0), mirzq(q));
            }
        }
        bool any_open = true;
        if (any_slow) {
            // the open fraction of step n blends each streamed link of a
            // cut cell with the bounce-back of its own opposite population
            mask0 = slow0 ? link_mask_e(c, e0, i, j, k) : 0x7fffeu;
            any_open = (mask0 & 0x7fffeu) != 0;
#pragma unroll
            for (int q = 1; q < 19; q++) {
                const int cxi = cxq(q), cyi = cyq(q), czi = czq(q);
                const bool yout = !yper && ((cyi < 0 && j == ny - 1) || (cyi > 0 && j == 0));
                const int sk = k - czi;
                if (yout || !(sk >= 0 && sk < nz)) continue;  // uniform
                const bool xout = (xlo && !xio && cxi > 0) || (xhi && cxi < 0);
                int lo;
                bool w;
                T frac = any_open ? link_geo_e(c, e0[c9q(q)], i, j, k, q, lo, w) : T(0);
                T bl = frac * f[q] + (T(1.0) - frac) * loc[oppq(q)];
                f[q] = (slow0 && !xout) ? bl : f[q];
            }
        }
        stream_edges(a, i, k, any_open, f);
        // (C) moving-wall IBB on near-obstacle cells (CODEX:3674-3712)
        const bool ibb_lane = slow0 && (mask0 & 0x7fffeu) != 0x7fffeu;
        if (__builtin_amdgcn_ballot_w64(ibb_lane) != 0) {
            const T* g = a.s0.ibb + col;
            T gw[19];
#pragma unroll
            for (int q = 1; q < 19; q++) gw[q] = g[(size_t)(q - 1) * plane];
#pragma unroll
            for (int q = 1; q < 19; q++) {
                int lo;
                bool w;
                T fo = link_geo_e(c, e0[c9q(q)], i, j, k, q, lo, w);
                if (!any_open) fo = T(0);
                T fw = loc[oppq(q)] - gw[q];
                T v = f[q] * fo + fw * (T(1.0) - fo);
                f[q] = ibb_lane ? v : f[q];
            }
        }
    }
#undef LD
    // update_bed_populations_and_reinit of step n+1 (CODEX:2894-2940)
    bool open0 = true;
    if (__builtin_amdgcn_ballot_w64(slow1) != 0) {
        T e1[9], fe[19];
#pragma unroll
        for (int e = 0; e < 9; e++) e1[e] = a.s1.eta9[(size_t)e * plane + col];
        const T* fs = a.s1.fsol + col;
#pragma unroll
        for (int q = 0; q < 19; q++) fe[q] = fs[(size_t)q * plane];
        const uint32_t mask = link_mask_e(c, e1, i, j, k);
        const bool op = (mask & 0x7fffeu) != 0;
        const bool cond = slow1 && (mask & 0x7fffeu) != 0x7fffeu;
        open0 = slow1 ? op : true;
#pragma unroll
        for (int q = 0; q < 19; q++) {
            int lo;
            bool w;
            T frac = link_geo_e(c, e1[c9q(q)], i, j, k, q, lo, w);
            if (q == 0) lo = op ? 1 : 0;
            else if (!op) frac = T(0);
            T v = frac * f[q] + (T(1.0) - frac) * fe[q];
            f[q] = (cond && !lo) ? v : f[q];
        }
    }
    open0_out = open0;
}

// Pull of step n into a cell above the wall band (every link open with
// fraction 1): the rules of band_gather without the bed, arranged so that every
// decision depending on y or z is uniform over the wave (a wave lies in one
// row of one plane: j is read from the first lane), and the rules of a
// non-periodic x boundary are a fix-up of the edge cells.
template <class S, class T>
__device__ __forceinline__ void lean_gather(const FusedArgs<S, T>& a, int i, int j_lane, int k, T f[19]) {
    const Consts<T>& c = a.c;
    const int j = __builtin_amdgcn_readfirstlane(j_lane);
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const uint32_t n = (uint32_t)nx * ny * nz;
    const uint32_t plane = (uint32_t)nx * ny;
    const uint32_t id = (uint32_t)k * plane + (uint32_t)j * nx + i;
    const Codec& cod = c.store;
    const S* fin = a.fin;
#define LD(q, idx) Store<S, T>::load(fin + (size_t)(q) * n, (uint32_t)(idx), (q), cod)
    if (a.mode == 0) {
#pragma unroll
        for (int q = 0; q < 19; q++) f[q] = LD(q, id);
        return;
    }
    const bool xper = c.x_boundary == LBM_XB_PERIODIC, xio = c.x_boundary == LBM_XB_INFLOW_OUTFLOW;
    const bool yper = c.y_boundary == LBM_YB_PERIODIC;
    const bool free_slip = c.top_boundary == LBM_TOP_FREE_SLIP;
    // source columns: periodic wrap, else the edge cell itself (the
    // outflow rule; the other edge rules are applied below)
    const int im = i == 0 ? (xper ? nx - 1 : 0) : i - 1;
    const int ip = i == nx - 1 ? (xper ? 0 : nx - 1) : i + 1;
    const int jm = j == 0 ? (yper ? ny - 1 : 0) : j - 1;
    const int jp = j == ny - 1 ? (yper ? 0 : ny - 1) : j + 1;
    f[0] = LD(0, id);
#pragma unroll
    for (int q = 1; q < 19; q++) {
        const int cxi = cxq(q), cyi = cyq(q), czi = czq(q);
        const int si = cxi > 0 ? im : (cxi < 0 ? ip : i);
        const int sj = cyi > 0 ? jm : (cyi < 0 ? jp : j);
        const bool yout = !yper && ((cyi < 0 && j == ny - 1) || (cyi > 0 && j == 0));
        const int sk = k - czi;
        T out;
        if (yout)
            out = LD(oppq(q), id);
        else if (sk >= 0 && sk < nz)
            out = LD(q, (uint32_t)sk * plane + (uint32_t)sj * nx + si);
        else if (sk >= nz && free_slip)
            out = LD(mirzq(q), (uint32_t)(nz - 1) * plane + (uint32_t)sj * nx + si);
        else
            out = LD(q, id);
        f[q] = out;
    }
    if (!xper && ((i == 0) || (xio ? false : i == nx - 1))) {
        // solid x walls: bounce back (x-outside source); inflow column:
        // equilibrium ghost at the inflow velocity (CODEX:3418-3551 + the
        // inflow/outflow rule); a y-outside source still bounces back
#pragma unroll
        for (int q = 1; q < 19; q++) {
            const int cxi = cxq(q), cyi = cyq(q), czi = czq(q);
            if (!((i == 0 && cxi > 0) || (i == nx - 1 && cxi < 0))) continue;
            const bool yout = !yper && ((cyi < 0 && j == ny - 1) || (cyi > 0 && j == 0));
            if (!xio || yout) {
                f[q] = LD(oppq(q), id);
                continue;
            }
            const int sj = cyi > 0 ? jm : (cyi < 0 ? jp : j);
            const int sk = k - czi;
            if (sk >= 0 && sk < nz)
                f[q] = feq(c.rho0, a.inflow[(size_t)sk * ny + sj], T(0), T(0), q);
            else if (sk >= nz && free_slip)
                f[q] = feq(c.rho0, a.inflow[(size_t)(nz - 1) * ny + sj], T(0), T(0), mirzq(q));
            else
                f[q] = LD(q, id);
        }
    }
#undef LD
    stream_edges(a, i, k, true, f);
}

// rho and u of step n+1 from the conditioned populations, in the order of
// the staged macro kernel (CODEX:2962-2987); the top drive fixes u on the
// top plane (CODEX:3650-3662)
template <class T>
__device__ __forceinline__ void fused_moments(const Consts<T>& c, int k, const T f[19], T& rr, T& un, T& vn, T& wn) {
    rr = T(0);
#pragma unroll
    for (int q = 0; q < 19; q++) rr += f[q];
    T jx = T(0), jy = T(0), jz = T(0);
#pragma unroll
    for (int q = 0; q < 19; q++) {
        if (cxq(q)) jx = (q == 1) ? f[q] : (cxq(q) > 0 ? jx + f[q] : jx - f[q]);
        if (cyq(q)) jy = (q == 3) ? f[q] : (cyq(q) > 0 ? jy + f[q] : jy - f[q]);
        if (czq(q)) jz = (q == 5) ? f[q] : (czq(q) > 0 ? jz + f[q] : jz - f[q]);
    }
    T inv = T(1.0) / rr;
    un = jx * inv;
    vn = jy * inv;
    wn = jz * inv;
    if (c.top_drive && k == c.nz - 1) {
        un = c.u_top_lb;
        vn = T(0);
        wn = T(0);
    }
}

// cell of this thread; false outside the grid
// XCD-aware order of the tiles of a plane.  The blocks of a launch are dealt
// to the 8 XCDs in turn (block L to XCD L mod 8) and each XCD has an L2 of
// its own.  With rows of tiles a multiple of 8 (so a plane is too), the b-th
// tile of a plane in dispatch order is taken from strip b mod 8, a band of
// gy / 8 consecutive rows of tiles: the velocity rows j +- 1 that a
// collision tile reads were written through the same L2, except at the 16
// strip edges.  Planes k +- 1 keep the same XCD either way.
__device__ __forceinline__ void tile_of(int b, int gx, int gy, int& bxi, int& byi) {
    const int s = b >> 3;
    byi = (b & 7) * (gy >> 3) + s / gx;
    bxi = s % gx;
}

template <class S, class T>
__device__ __forceinline__ bool fused_cell(const FusedArgs<S, T>& a, int& i, int& j, int& k) {
    int bxi = (int)blockIdx.x, byi = (int)blockIdx.y;
    if (a.strips) tile_of(bxi + (int)gridDim.x * byi, (int)gridDim.x, (int)gridDim.y, bxi, byi);
    i = bxi * (int)blockDim.x + (int)threadIdx.x;
    j = byi * (int)blockDim.y + (int)threadIdx.y;
    k = blockIdx.z + a.plane0;
    return i < a.c.nx && j < a.c.ny;
}

// waves per SIMD the kernels are compiled for (register budget 512 / waves):
// in FP32 the lean moments pass fits 8 waves, the lean collision 5, the band
// passes 4; FP64 arithmetic halves each
template <class T, bool BAND>
struct FusedOcc {
    static constexpr int k1 = (sizeof(T) == 8 ? 4 : 8) / (BAND ? 2 : 1);
    static constexpr int k2 = sizeof(T) == 8 ? 2 : (BAND ? 4 : 5);
};

// velocity / open-flag exchange between the passes: plain accesses between
// two launches, device-coherent ones (agent scope, no cache flush) inside
// the wavefront launch, where the reader may run on another XCD
template <bool COH, class V>
__device__ __forceinline__ void xput(V* p, V v) {
    if (COH)
        __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else
        *p = v;
}
template <bool COH, class V>
__device__ __forceinline__ V xget(const V* p) {
    if (COH) return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return *p;
}

template <class S, class T, bool BAND, bool COH = false>
__device__ __forceinline__ void moments_cell(const FusedArgs<S, T>& a, int i, int j, int k) {
    const Consts<T>& c = a.c;
    T f[19];
    bool open0 = true;
    if (BAND)
        band_gather(a, i, j, k, f, open0);
    else
        lean_gather(a, i, j, k, f);
    T rr, un, vn, wn;
    fused_moments(c, k, f, rr, un, vn, wn);
    const uint32_t id = ((uint32_t)k * c.ny + j) * c.nx + i;
    xput<COH>(a.ux + id, un);
    xput<COH>(a.uy + id, vn);
    xput<COH>(a.uz + id, wn);
    if (BAND) xput<COH>(a.open + id, (uint8_t)(open0 ? 1 : 0));
}

template <class S, class T, bool BAND>
__global__ __launch_bounds__(FB_THREADS, (FusedOcc<T, BAND>::k1)) void k_moments(FusedArgs<S, T> a) {
    int i, j, k;
    if (!fused_cell(a, i, j, k)) return;
    moments_cell<S, T, BAND>(a, i, j, k);
}

// LES viscosity (CODEX:3011-3141) and collision (CODEX:3143-3416, air mode)
// of one cell whose conditioned populations, moments and neighbour
// velocities are known, in place; returns omega.  nbu(comp, dir) returns the
// neighbour velocity component comp (0..2) in direction dir = x+, x-, y+, y-,
// z+, z- (x and y wrap periodically, CODEX:3047-3053; z clamps at the ends).
// collision of a launch: any model (chosen per cell from the constants), or
// one fixed model (no sponge, no backscatter) so the kernel holds only its code
// blocks of 256 threads per CU the wavefront step is compiled for (6: 80
// VGPRs and 24 B per lane of scratch; 5, 84 VGPRs without scratch, ran 3 %
// slower on the 256x128x128 waves step)
constexpr int WAVE_OCC = 6;

template <int CM, class S, class T, class NB>
__device__ __forceinline__ T les_collide(const FusedArgs<S, T>& a, int i, int j, int k, uint32_t id, T f[19], bool open_cur, bool open_below,
                                         T rcur, T ucur, T vcur, T wcur, const NB& nbu) {
    const Consts<T>& c = a.c;
    T om = T(1.0) / c.tau0;
    if (c.use_les) {
        if (open_cur) {
            const bool near_wall = k > 0 && !open_below;
            T phi = T(0), uwx = T(0), uwy = T(0);
            if (near_wall) {
                phi = cell_phi(c, a.s1, i, j, k);
                if (c.include_bed) {
                    const size_t col = (size_t)j * c.nx + i;
                    uwx = a.s1.uw[col];
                    uwy = a.s1.vw[col];
                }
            }
            om = les_omega(c, nbu(0, 0), nbu(0, 1), nbu(0, 2), nbu(0, 3), nbu(0, 4), nbu(0, 5), nbu(1, 0), nbu(1, 1), nbu(1, 2),
                           nbu(1, 3), nbu(1, 4), nbu(1, 5), nbu(2, 0), nbu(2, 1), nbu(2, 2), nbu(2, 3), nbu(2, 4), nbu(2, 5), near_wall,
                           phi, ucur, vcur, uwx, uwy, (T*)nullptr, (T*)nullptr);
            if (k <= c.omega_band_top) a.omega[id] = om;
        } else if (k <= c.omega_band_top) {
            om = a.omega[id];  // inactive cells keep their last omega
        }
    }
    if (CM == CM_REG) {
        fused_regularized(f, rcur, ucur, vcur, wcur, om);
    } else if (CM == CM_KBC) {
        fused_kbc(f, rcur, ucur, vcur, wcur, om);
    } else if (c.collision_reg) {
        fused_regularized(f, rcur, ucur, vcur, wcur, om);
    } else if (c.top_sponge && k > c.nz - 3) {
        sponge_relax(f, rcur, ucur, vcur, wcur);
    } else if (c.c_backscatter > T(0)) {
        const bool bs = open_cur;
        const float xi = bs ? bs_uniform(c.seed, a.step, id) : 0.0f;
        collide_kbc<T, false>(f, rcur, ucur, vcur, wcur, om, T(0), bs, c.c_backscatter, xi);
    } else {
        fused_kbc(f, rcur, ucur, vcur, wcur, om);
    }
    return om;
}

// les_collide, then the store of f* (and of rho, omega on the last step)
template <int CM, class S, class T, class NB>
__device__ __forceinline__ void les_collide_store(const FusedArgs<S, T>& a, int i, int j, int k, uint32_t id, T f[19], bool open_cur,
                                                  bool open_below, T rcur, T ucur, T vcur, T wcur, const NB& nbu) {
    const Consts<T>& c = a.c;
    const uint32_t n = (uint32_t)c.nx * c.ny * c.nz;
    const T om = les_collide<CM>(a, i, j, k, id, f, open_cur, open_below, rcur, ucur, vcur, wcur, nbu);
    S* const fout = a.fout;
    const uint32_t dbase = c.store.stochastic ? dither_base(a.step, id) : 0u;
#pragma unroll
    for (int q = 0; q < 19; q++) Store<S, T>::store(fout + (size_t)q * n, id, q, f[q], c.store, dbase);
    if (a.write_macro) {
        a.rho[id] = rcur;
        a.omega[id] = om;
    }
}

template <class S, class T, bool BAND, int CM, bool COH = false>
__device__ __forceinline__ void collide_cell(const FusedArgs<S, T>& a, int i, int j, int k) {
    const Consts<T>& c = a.c;
    const int nx = c.nx, ny = c.ny, nz = c.nz;
    const uint32_t plane = (uint32_t)nx * ny;
    const uint32_t id = (uint32_t)k * plane + (uint32_t)j * nx + i;
    T f[19];
    bool open_cur = true;
    if (BAND)
        band_gather(a, i, j, k, f, open_cur);
    else
        lean_gather(a, i, j, k, f);
    T rcur, ucur, vcur, wcur;
    fused_moments(c, k, f, rcur, ucur, vcur, wcur);
    const bool open_below = BAND && k > 0 ? xget<COH>(a.open + (id - plane)) != 0 : true;
    const int ip = i == nx - 1 ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
    const int jp = j == ny - 1 ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
    const uint32_t row = (uint32_t)k * plane + (uint32_t)j * nx;
    const uint32_t nbi[6] = {row + ip, row + im, (uint32_t)k * plane + (uint32_t)jp * nx + i,
                             (uint32_t)k * plane + (uint32_t)jm * nx + i, (k + 1 < nz ? id + plane : id), (k > 0 ? id - plane : id)};
    const T* u3[3] = {a.ux, a.uy, a.uz};
    auto nbu = [&](int comp, int dir) { return xget<COH>(u3[comp] + nbi[dir]); };
    les_collide_store<CM>(a, i, j, k, id, f, open_cur, open_below, rcur, ucur, vcur, wcur, nbu);
}

template <class S, class T, bool BAND, int CM>
__global__ __launch_bounds__(FB_THREADS, (FusedOcc<T, BAND>::k2)) void k_collide(FusedArgs<S, T> a) {
    int i, j, k;
    if (!fused_cell(a, i, j, k)) return;
    collide_cell<S, T, BAND, CM>(a, i, j, k);
}

// both collision forms in one launch over every plane: the branch on k is
// uniform over a block, so each block runs one form, and the step saves the
// boundary between two launches (the two forms need about the same registers)
template <class S, class T, int CM>
__global__ __launch_bounds__(FB_THREADS, 6) void k_collide_all(FusedArgs<S, T> a) {
    int i, j, k;
    if (!fused_cell(a, i, j, k)) return;
    if (k < a.band_planes)
        collide_cell<S, T, true, CM>(a, i, j, k);
    else
        collide_cell<S, T, false, CM>(a, i, j, k);
}

// ---------------------------------------------------------------------------
// Wavefront form of the fused step: both passes in one launch
// ---------------------------------------------------------------------------
//
// One launch runs both passes.  Block t is one block-sized tile (FB_THREADS
// cells of one plane) of one pass, and the block indices run through the
// planes in the order
//   M(0) .. M(L-1),  M(L) C(0),  M(L+1) C(1),  ...,  C(nz-L) .. C(nz-1)
// (M: moments pass, C: collision pass, L: lag in planes).  The collision of
// plane k thus re-reads the populations that the moments of planes
// k-1 .. k+1 read a few planes earlier, while they are still held in the L2
// or the memory-side cache, and the step has no launch boundary between the
// passes.
//
// A collision tile needs the velocities (and, on band planes, the open
// flags) of the moments tiles around it: the 3 x 3 tiles of its own plane
// and the tile above and below.  Each moments tile publishes with a flag of
// its own (no read-modify-write): its stores of u and of the flag are
// device-coherent (agent scope, no cache flush), and the flag is written
// once the stores are acknowledged.  The flag holds the launch's epoch, so
// no launch needs a reset.  A collision tile polls its at most 11 flags.
// Every tile it waits on has a lower block index; the hardware dispatches
// the blocks of each XCD in increasing index order, so the lowest
// unfinished block is always resident and its waits end.  Each wait is also
// bounded: one that runs out sets an error flag (reported by the host) and
// every later wait of the launch gives up at once.
struct WaveQ {
    int* flag;   // [nz * bpp] epoch of the last launch whose moments tile is done
    int* err;    // [0] a wait ran out
    int epoch;   // this launch (>= 1)
    int lag;     // planes of moments ahead of the collision
    int bpp;     // tiles per plane
    int gx, gy;  // tiles per row of tiles (x), rows of tiles (y)
    int bx;      // tile extent in x (64, 128 or 256)
};
constexpr int WAVE_SPIN_LIMIT = 1 << 22;  // s_sleep(2) polls: a few hundred milliseconds

// pass (0 moments, 1 collision) and plane of the u-th plane unit
__device__ __forceinline__ void wave_unit(int u, int L, int nz, int& pass, int& k) {
    if (u < L) {
        pass = 0;
        k = u;
        return;
    }
    int v = u - L;
    const int npair = nz - L;
    if (v < 2 * npair) {
        pass = v & 1;
        k = (v >> 1) + (pass ? 0 : L);
        return;
    }
    pass = 1;
    k = npair + (v - 2 * npair);
}

// flag index polled by lane l (0 .. 10) of a collision tile, -1 for none:
// lanes 0..8 the 3 x 3 tiles of plane k, 9 and 10 the tiles below / above
__device__ __forceinline__ int wave_dep(const WaveQ& w, int nz, int k, int bxi, int byi, int l) {
    int kk = k, tx = bxi, ty = byi;
    if (l < 9) {
        const int dx = l % 3 - 1, dy = l / 3 - 1;
        tx += dx;
        ty += dy;
        // the LES differences wrap in x and y whatever the boundaries
        // (CODEX:3047-3053)
        if (tx < 0) tx = w.gx - 1;
        if (tx >= w.gx) tx = 0;
        if (ty < 0) ty = w.gy - 1;
        if (ty >= w.gy) ty = 0;
    } else if (l < 11) {
        kk = k + (l == 9 ? -1 : 1);
        if (kk < 0 || kk >= nz) return -1;
    } else {
        return -1;
    }
    return kk * w.bpp + ty * w.gx + tx;
}

// tile t of a case's wavefront step (the body of k_step_wave and of the
// batched form)
template <class S, class T, int CM>
__device__ __forceinline__ void wave_tile(const FusedArgs<S, T>& a, const WaveQ& w, int t) {
    const int tid = (int)threadIdx.x;
    const int nz = a.c.nz;
    const int u = t / w.bpp, slot = t - u * w.bpp;
    int pass, k;
    wave_unit(u, w.lag, nz, pass, k);
    int bxi, byi;
    if (a.strips) {
        tile_of(slot, w.gx, w.gy, bxi, byi);
    } else {
        byi = slot / w.gx;
        bxi = slot - byi * w.gx;
    }
    const int b = byi * w.gx + bxi;  // the tile's own index (its flag)
    const int i = bxi * w.bx + tid % w.bx, j = byi * (FB_THREADS / w.bx) + tid / w.bx;
    const bool in = i < a.c.nx && j < a.c.ny;
    const bool band = k < a.band_planes;
    if (pass == 0) {
        if (in) {
            if (band)
                moments_cell<S, T, true, true>(a, i, j, k);
            else
                moments_cell<S, T, false, true>(a, i, j, k);
        }
        // publish: every wave waits until its device-coherent stores are
        // acknowledged, then one thread writes the tile's flag
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&w.flag[k * w.bpp + b], w.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        if (tid < 64) {
            const int d = wave_dep(w, nz, k, bxi, byi, tid);
            bool wait = d >= 0;
            int spins = 0;
            while (__builtin_amdgcn_ballot_w64(wait) != 0) {
                if (wait) wait = __hip_atomic_load(&w.flag[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != w.epoch;
                if (__builtin_amdgcn_ballot_w64(wait) == 0) break;
                if (++spins > WAVE_SPIN_LIMIT || __hip_atomic_load(w.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    if (tid == 0) __hip_atomic_store(w.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        // the velocity and flag loads after this barrier are device-coherent
        __syncthreads();
        if (in) {
            if (band)
                collide_cell<S, T, true, CM, true>(a, i, j, k);
            else
                collide_cell<S, T, false, CM, true>(a, i, j, k);
        }
    }
}

template <class S, class T, int CM>
__global__ __launch_bounds__(FB_THREADS, WAVE_OCC) void k_step_wave(FusedArgs<S, T> a, WaveQ w) {
    wave_tile<S, T, CM>(a, w, (int)blockIdx.x);
}

// ---------------------------------------------------------------------------
// Tile-resident form of the fused step: one persistent, co-resident grid
// ---------------------------------------------------------------------------
//
// Block b of a grid of R blocks runs the tiles b, b + R, b + 2R, ... in
// plane-major order, one iteration ahead: in iteration i it runs the
// moments pass of its next tile (pull, conditioning, u; the pulled
// populations go to one of two LDS buffers), publishes that tile's flag,
// waits for the flags of the tiles whose velocities the collision of its
// current tile needs (the 3 x 3 tiles of its plane, the tiles above and
// below), and runs that collision with the populations it pulled one
// iteration earlier, read back from LDS: the collision's second pull of the
// wavefront form (76 B per update in FP32) is gone.  The waits point
// forward (the tile above is bpp tiles later), so the grid is launched
// cooperatively (all blocks resident) and only when R exceeds the longest
// forward wait D (bpp + gx + 1 tiles in plane order; 2 bpp with the XCD
// strip order of the tiles of a plane).  Then the lowest unfinished
// collision, of tile t0, waits only on the moments of tiles t0 < t <= t0 + D;
// the block of such a tile runs its moments at the start of an iteration
// whose earlier collisions lie R or more tiles before t (below t0, so
// finished), and moments never wait, so t0 completes (the host checks the
// size; the waits are bounded as in k_step_wave all the same).
// tiles of RES_THREADS cells, RES_OCC waves per SIMD (the launch bound:
// 128 VGPRs; 2 x 19.5 KB of LDS per block, 4 blocks per CU).  Tiles of 128
// cells (8 blocks per CU, 134 VGPRs) ran 20 % slower.
constexpr int RES_THREADS = 256;
constexpr int RES_OCC = 4;
struct ResTile {
    int k, bxi, byi, b, i, j;
    bool in, band;
    uint32_t id;
};
template <class S, class T, int CM>
__global__ __launch_bounds__(RES_THREADS, RES_OCC) void k_step_resident(FusedArgs<S, T> a, WaveQ w, int total) {
    __shared__ T sf[2][19][RES_THREADS];
    __shared__ uint8_t sopen[2][RES_THREADS];
    const int tid = (int)threadIdx.x;
    const int nx = a.c.nx, ny = a.c.ny, nz = a.c.nz;
    const uint32_t plane = (uint32_t)nx * ny;
    auto tile = [&](int t) {
        ResTile x;
        x.k = t / w.bpp;
        const int slot = t - x.k * w.bpp;
        if (a.strips) {
            tile_of(slot, w.gx, w.gy, x.bxi, x.byi);
        } else {
            x.byi = slot / w.gx;
            x.bxi = slot - x.byi * w.gx;
        }
        x.b = x.byi * w.gx + x.bxi;  // the tile's own index (its flag)
        x.i = x.bxi * w.bx + tid % w.bx;
        x.j = x.byi * (RES_THREADS / w.bx) + tid / w.bx;
        x.in = x.i < nx && x.j < ny;
        x.band = x.k < a.band_planes;
        x.id = (uint32_t)x.k * plane + (uint32_t)x.j * nx + x.i;
        return x;
    };
    // the moments pass of tile x into LDS buffer buf, then its flag
    auto moments = [&](const ResTile& x, int buf) {
        if (x.in) {
            T f[19];
            bool open_cur = true;
            if (x.band)
                band_gather(a, x.i, x.j, x.k, f, open_cur);
            else
                lean_gather(a, x.i, x.j, x.k, f);
            T r, u, v, ww;
            fused_moments(a.c, x.k, f, r, u, v, ww);
            xput<true>(a.ux + x.id, u);
            xput<true>(a.uy + x.id, v);
            xput<true>(a.uz + x.id, ww);
            if (x.band) xput<true>(a.open + x.id, (uint8_t)(open_cur ? 1 : 0));
#pragma unroll
            for (int q = 0; q < 19; q++) sf[buf][q][tid] = f[q];
            sopen[buf][tid] = open_cur ? 1 : 0;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) __hip_atomic_store(&w.flag[x.k * w.bpp + x.b], w.epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
    int t = (int)blockIdx.x;
    if (t >= total) return;
    ResTile cur = tile(t);
    moments(cur, 0);
    for (int it = 0;; it++) {
        const int tn = t + (int)gridDim.x;
        const bool more = tn < total;
        ResTile nxt = cur;
        if (more) {
            nxt = tile(tn);
            moments(nxt, (it + 1) & 1);
        }
        if (tid < 64) {
            const int d = wave_dep(w, nz, cur.k, cur.bxi, cur.byi, tid);
            bool wait = d >= 0;
            int spins = 0;
            while (__builtin_amdgcn_ballot_w64(wait) != 0) {
                if (wait) wait = __hip_atomic_load(&w.flag[d], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != w.epoch;
                if (__builtin_amdgcn_ballot_w64(wait) == 0) break;
                if (++spins > WAVE_SPIN_LIMIT || __hip_atomic_load(w.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
                    if (tid == 0) __hip_atomic_store(w.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                __builtin_amdgcn_s_sleep(2);
            }
        }
        __syncthreads();  // the velocity and flag loads below are device-coherent
        if (cur.in) {
            const int i = cur.i, j = cur.j, k = cur.k;
            const uint32_t id = cur.id;
            T f[19];
#pragma unroll
            for (int q = 0; q < 19; q++) f[q] = sf[it & 1][q][tid];
            const bool open_cur = sopen[it & 1][tid] != 0;
            T rcur, ucur, vcur, wcur;
            fused_moments(a.c, k, f, rcur, ucur, vcur, wcur);  // as in the moments pass, bit for bit
            const bool open_below = cur.band && k > 0 ? xget<true>(a.open + (id - plane)) != 0 : true;
            const int ip = i == nx - 1 ? 0 : i + 1, im = i == 0 ? nx - 1 : i - 1;
            const int jp = j == ny - 1 ? 0 : j + 1, jm = j == 0 ? ny - 1 : j - 1;
            const uint32_t row = (uint32_t)k * plane + (uint32_t)j * nx;
            const uint32_t nbi[6] = {row + ip, row + im, (uint32_t)k * plane + (uint32_t)jp * nx + i,
                                     (uint32_t)k * plane + (uint32_t)jm * nx + i, (k + 1 < nz ? id + plane : id), (k > 0 ? id - plane : id)};
            const T* u3[3] = {a.ux, a.uy, a.uz};
            auto nbu = [&](int comp, int dir) { return xget<true>(u3[comp] + nbi[dir]); };
            les_collide_store<CM>(a, i, j, k, id, f, open_cur, open_below, rcur, ucur, vcur, wcur, nbu);
        }
        if (!more) break;
        __syncthreads();  // buffer it & 1 is the next moments pass's but one
        t = tn;
        cur = nxt;
    }
}

// Several cases advanced by one launch (the ensemble server's batches).  The
// first half of the grid holds the moments tiles of every case (case c:
// blocks block0[c] .. block0[c] + nz bpp - 1, plane by plane), the second
// half their collision tiles in the same order, so every tile a collision
// tile waits on has a lower block index and has usually finished long
// before (the cases' tiles run side by side, with no lag to tune).  Each
// case's flags use a lag of nz planes, which is this order.  The table holds
// each case's arguments of this step.
template <class S, class T>
struct WaveCase {
    FusedArgs<S, T> a;
    WaveQ w;     // lag = nz
    int block0;  // first moments block of the case (the half of the grid: + half)
    int pad;
};
template <class S, class T, int CM>
__global__ __launch_bounds__(FB_THREADS, 6) void k_step_wave_batch(const WaveCase<S, T>* __restrict__ tab, int ncases, int half) {
    int t = (int)blockIdx.x;
    const bool coll = t >= half;
    if (coll) t -= half;
    int lo = 0, hi = ncases - 1;  // the last case whose block0 <= t
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tab[mid].block0 <= t)
            lo = mid;
        else
            hi = mid - 1;
    }
    const FusedArgs<S, T> a = tab[lo].a;
    const WaveQ w = tab[lo].w;
    const int local = t - tab[lo].block0;
    wave_tile<S, T, CM>(a, w, coll ? a.c.nz * w.bpp + local : local);
}

// ===========================================================================
// reductions (FP64)
// ===========================================================================
template <class S, class T>
struct RedArgs {
    Consts<T> c;
    Slot<T> s;        // time level of the state (pending stream of a post-collision state)
    const S* f;       // populations: post-stream (SRC 0) or post-collision (SRC 1)
    const T *ux, *uy, *uz, *mass, *fill, *u_inlet, *inflow, *fs_h;
    const uint8_t* flags;
    const int* type;
};

// one cell's contribution to reduction `what`; SRC 1 streams a
// post-collision state on the fly (the post-stream populations the
// materialised download returns, without a buffer for them)
template <int SRC, class S, class T>
__device__ __forceinline__ double reduce_cell(const RedArgs<S, T>& a, int what, size_t id) {
    const Consts<T>& c = a.c;
    const size_t n = (size_t)c.nx * c.ny * c.nz;
    if (what == LBM_REDUCE_VOF_MASS) return (double)a.mass[id];
    if (what == LBM_REDUCE_VOF_VOLUME) return (double)a.fill[id];
    if (what == LBM_REDUCE_MAX_SPEED) {
        bool act = a.flags[id] & 1u;
        if (c.fs_water && a.type[id] == LBM_FS_GAS) act = false;
        if (!act) return 0.0;
        const double x = a.ux[id], y = a.uy[id], z = a.uz[id];
        const double v = sqrt(x * x + y * y + z * z);
        return v != v ? 0.0 : v;  // np.nanmax ignores NaN
    }
    if (what == LBM_REDUCE_MASS && !(a.flags[id] & 1u)) return 0.0;
    T fo[19];
    if (SRC == 1) {
        int ty, k, n1, n2, n3;
        stream_bc_cell<S, T>(c, a.s, a.f, a.ux, a.uy, a.uz, nullptr, a.type, a.fill, a.u_inlet, a.inflow, a.fs_h, a.flags, id, fo, ty, k, n1,
                             n2, n3);
    } else {
#pragma unroll
        for (int q = 0; q < 19; q++) fo[q] = Store<S, T>::load(a.f, (size_t)q * n + id, q, c.store);
    }
    double v = 0.0;
    if (what == LBM_REDUCE_MASS) {
#pragma unroll
        for (int q = 0; q < 19; q++) v += (double)fo[q];
    } else {
#pragma unroll
        for (int q = 0; q < 19; q++)
            if (!__builtin_isfinite(fo[q])) v += 1.0;
    }
    return v;
}

template <int SRC, class S, class T>
__global__ __launch_bounds__(RED_THREADS) void k_reduce_part(RedArgs<S, T> a, int what, double* part) {
    const size_t n = (size_t)a.c.nx * a.c.ny * a.c.nz;
    const size_t base = (size_t)blockIdx.x * RED_CHUNK + threadIdx.x;
    const bool is_max = what == LBM_REDUCE_MAX_SPEED;
    double v = 0.0;
    for (int r = 0; r < RED_PER_THREAD; r++) {
        const size_t id = base + (size_t)r * RED_THREADS;
        if (id < n) {
            const double x = reduce_cell<SRC>(a, what, id);
            v = is_max ? fmax(v, x) : v + x;
        }
    }
    v = is_max ? block_tree<true>(v) : block_tree<false>(v);
    if (threadIdx.x == 0) part[blockIdx.x] = v;
}

// partials -> out[0]; `stride` and `count` select one of several
// interleaved partial sets
__global__ __launch_bounds__(RED_THREADS) void k_reduce_final(const double* part, int count, int stride, int is_max, double* out) {
    double v = 0.0;
    for (int i = (int)threadIdx.x; i < count; i += RED_THREADS) {
        const double x = part[(size_t)i * stride];
        v = is_max ? fmax(v, x) : v + x;
    }
    v = is_max ? block_tree<true>(v) : block_tree<false>(v);
    if (threadIdx.x == 0) out[0] = v;
}

template <class S, class T>
__global__ __launch_bounds__(NT) void k_store_populations(Consts<T> c, const T* src, S* dst) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (idx >= 19 * n) return;
    int q = (int)(idx / n);
    Store<S, T>::store(dst, idx, q, src[idx], c.store, dither_base(c.store_step, (uint32_t)(idx - (size_t)q * n)));
}

template <class S, class T>
__global__ __launch_bounds__(NT) void k_load_populations(Consts<T> c, const S* src, T* dst) {
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (idx >= 19 * n) return;
    int q = (int)(idx / n);
    dst[idx] = Store<S, T>::load(src, idx, q, c.store);
}

// ===========================================================================
// launchers
// ===========================================================================
// air mode, the bed the only obstacle and its gap the open fraction (the band
// passes evaluate the bed's link geometry only)
bool fused_supported(const lbm_config_t& c) {
    return c.physics_mode == LBM_PHYS_AIR && !c.fs_enabled && !c.include_cube && !c.phi_fraction && c.nz >= 3;
}

// compute_phi_slope_corrected (CODEX:2432-2490) into a 3-D field, for the
// boundary-geometry modes that sample it at the link endpoints
template <class T>
__global__ __launch_bounds__(NT) void k_phi(Consts<T> c, Slot<T> s, T* phi) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    phi[id] = cell_phi(c, s, i, j, k);
}

template <class T>
static void bed2d_t(const lbm_config_t& cfg, DevBufs& b, float t, int slot, hipStream_t s) {
    Consts<T> c = make_consts<T>(cfg, -1);
    size_t n2 = b.n2;
    const int prev = (slot + b.ring - 1) % b.ring;
    k_bed2d_a<T><<<nblocks(n2, 64), 64, 0, s>>>(c, (T)t, (T*)b.h[slot], (T*)b.uw[slot], (T*)b.vw[slot], (T*)b.ww[slot],
                                             (const T*)b.h[prev], (const T*)b.piecewise);
    k_bed2d_b<T><<<nblocks(n2, BED_COLS), 9 * BED_COLS, 0, s>>>(c, (const T*)b.h[slot], (T*)b.eta9[slot], (T*)b.denom[slot], b.kband[slot],
                                             (const T*)b.uw[slot], (const T*)b.vw[slot], (const T*)b.ww[slot], (T*)b.uw9[slot],
                                             b.ksolid[slot], (T*)b.fsol[slot], (T*)b.ibb[slot]);
    if (cfg.phi_fraction) k_phi<T><<<nblocks(b.n), NT, 0, s>>>(c, make_slot<T>(b, slot), (T*)b.phi);
    HIP_OK(hipGetLastError());
}
void launch_bed2d(const lbm_config_t& c, DevBufs& b, float t, int slot, hipStream_t s) {
    if (b.storage == LBM_STORE_FP64)
        bed2d_t<double>(c, b, t, slot, s);
    else
        bed2d_t<float>(c, b, t, slot, s);
}

// the time of step `step` as the step loop forms it (CODEX:3874)
static float step_time(const lbm_config_t& c, int64_t step) { return (float)((double)(step + 1) * c.dt_phys); }

template <class T>
static BedBatch<T> make_bed_batch(const lbm_config_t& cfg, DevBufs& b, int64_t step0, int nb) {
    if (nb < 1 || nb > BED_BATCH || b.ring != BED_RING) throw std::runtime_error("bed batch: bad size or ring");
    if (cfg.phi_fraction) throw std::runtime_error("bed batch: the phi modes keep one phi field per step");
    BedBatch<T> a{};
    for (int k = 0; k < nb; k++) {
        const int64_t st = step0 + k;
        const int sl = (int)(st % b.ring);
        a.t[k] = (T)step_time(cfg, st);
        a.tprev[k] = (T)step_time(cfg, st - 1);
        a.h[k] = (T*)b.h[sl];
        a.uw[k] = (T*)b.uw[sl];
        a.vw[k] = (T*)b.vw[sl];
        a.ww[k] = (T*)b.ww[sl];
        a.eta9[k] = (T*)b.eta9[sl];
        a.denom[k] = (T*)b.denom[sl];
        a.kband[k] = b.kband[sl];
        a.uw9[k] = (T*)b.uw9[sl];
        a.ksolid[k] = b.ksolid[sl];
        a.fsol[k] = (T*)b.fsol[sl];
        a.ibb[k] = (T*)b.ibb[sl];
    }
    a.h_last0 = (const T*)b.h[(int)((step0 - 1 + b.ring) % b.ring)];
    return a;
}
template <class T>
static void bed2d_batch_t(const lbm_config_t& cfg, DevBufs& b, int64_t step0, int nb, hipStream_t s) {
    Consts<T> c = make_consts<T>(cfg, -1);
    BedBatch<T> a = make_bed_batch<T>(cfg, b, step0, nb);
    k_bed2d_a_batch<T><<<dim3(nblocks(b.n2, 64), nb), 64, 0, s>>>(c, a, (const T*)b.piecewise);
    k_bed2d_b_batch<T><<<dim3(nblocks(b.n2, BED_COLS), nb), 9 * BED_COLS, 0, s>>>(c, a);
    HIP_OK(hipGetLastError());
}
template <class T>
static void fill_bed_case_t(const lbm_config_t& cfg, DevBufs& b, int64_t step0, int nb, int a0, int b0, void* out, int* na, int* nbb) {
    BedCase<T> e;
    memset(&e, 0, sizeof e);
    e.c = make_consts<T>(cfg, -1);
    e.a = make_bed_batch<T>(cfg, b, step0, nb);
    e.piecewise = (const T*)b.piecewise;
    e.nb = nb;
    e.a0 = a0;
    e.b0 = b0;
    memcpy(out, &e, sizeof e);
    *na = (int)nblocks(b.n2, 64) * nb;
    *nbb = (int)nblocks(b.n2, BED_COLS) * nb;
}
size_t bed_case_bytes(int storage) { return storage == LBM_STORE_FP64 ? sizeof(BedCase<double>) : sizeof(BedCase<float>); }
void fill_bed_case(const lbm_config_t& c, DevBufs& b, int64_t step0, int nb, int a0, int b0, void* out, int* na, int* nbb) {
    if (b.storage == LBM_STORE_FP64)
        fill_bed_case_t<double>(c, b, step0, nb, a0, b0, out, na, nbb);
    else
        fill_bed_case_t<float>(c, b, step0, nb, a0, b0, out, na, nbb);
}
void launch_bed_multi(int storage, const void* dev_table, int ncases, int na, int nbb, hipStream_t s) {
    if (storage == LBM_STORE_FP64) {
        k_bed2d_a_multi<double><<<na, 64, 0, s>>>((const BedCase<double>*)dev_table, ncases);
        k_bed2d_b_multi<double><<<nbb, 9 * BED_COLS, 0, s>>>((const BedCase<double>*)dev_table, ncases);
    } else {
        k_bed2d_a_multi<float><<<na, 64, 0, s>>>((const BedCase<float>*)dev_table, ncases);
        k_bed2d_b_multi<float><<<nbb, 9 * BED_COLS, 0, s>>>((const BedCase<float>*)dev_table, ncases);
    }
    HIP_OK(hipGetLastError());
}

void launch_bed2d_batch(const lbm_config_t& c, DevBufs& b, int64_t step0, int nb, hipStream_t s) {
    if (b.storage == LBM_STORE_FP64)
        bed2d_batch_t<double>(c, b, step0, nb, s);
    else
        bed2d_batch_t<float>(c, b, step0, nb, s);
}

template <class T>
static void geometry_fields_t(const lbm_config_t& cfg, DevBufs& b, int slot, int32_t* lo, void* lf, int32_t* wt, int32_t* near,
                              void* phi, hipStream_t s) {
    Consts<T> c = make_consts<T>(cfg, -1);
    k_geometry_fields<T><<<nblocks(b.n), NT, 0, s>>>(c, make_slot<T>(b, slot), lo, (T*)lf, wt, near, (T*)phi);
    HIP_OK(hipGetLastError());
}
void launch_geometry_fields(const lbm_config_t& c, DevBufs& b, int slot, int32_t* lo, void* lf, int32_t* wt, int32_t* near,
                            void* phi, hipStream_t s) {
    if (b.storage == LBM_STORE_FP64)
        geometry_fields_t<double>(c, b, slot, lo, lf, wt, near, phi, s);
    else
        geometry_fields_t<float>(c, b, slot, lo, lf, wt, near, phi, s);
}

template <class T>
static void cellflags_t(const lbm_config_t& cfg, DevBufs& b, int slot, hipStream_t s) {
    Consts<T> c = make_consts<T>(cfg, -1);
    k_cellflags<T><<<nblocks(b.n), NT, 0, s>>>(c, make_slot<T>(b, slot), b.cellflags);
    HIP_OK(hipGetLastError());
}
void launch_cellflags(const lbm_config_t& c, DevBufs& b, int slot, hipStream_t s) {
    if (b.storage == LBM_STORE_FP64)
        cellflags_t<double>(c, b, slot, s);
    else
        cellflags_t<float>(c, b, slot, s);
}

template <class T>
static VofArgs<T> make_vof(const lbm_config_t& cfg, DevBufs& b) {
    VofArgs<T> a;
    a.c = make_consts<T>(cfg, -1);
    a.rho = (T*)b.rho;
    a.ux = (T*)b.ux;
    a.uy = (T*)b.uy;
    a.uz = (T*)b.uz;
    a.fill = (T*)b.fill;
    a.mass = (T*)b.mass;
    a.mass2 = (T*)b.mass2;
    a.mdelta = (T*)b.mdelta;
    a.excess = (T*)b.excess;
    a.excess2 = (T*)b.excess2;
    a.sa = (T*)b.sa;
    a.sb = (T*)b.sb;
    a.sc = (T*)b.sc;
    a.sd = (T*)b.sd;
    a.se = (T*)b.se;
    a.smv = (T*)b.smv;
    a.type = b.type;
    a.type_prev = b.type_prev;
    a.excess_dir = b.excess_dir;
    a.ext_gas = b.ext_gas;
    a.si = b.si;
    a.mdirty = b.mdirty;
    a.rowdon = b.rowdon;
    a.act = b.rowact;
    a.act_prev = b.rowact + (size_t)cfg.nz * cfg.ny;
    a.iso = b.rowact + 2 * (size_t)cfg.nz * cfg.ny;
    a.band = 0;
    a.flags = b.cellflags;
    a.icnt = b.icnt;
    a.fs_h = (const T*)b.fs_h;
    a.dcnt = b.dcnt;
    return a;
}

// initialize_free_surface_boundary_populations (CODEX:2829-2875): fout is a
// copy of fin except, in open wet cells, the populations streamed from an
// open gas cell (or from above the lid), which take the gas-pressure closure
// of the pull stream feq_q(rho_fs, u) + feq_qbar(rho_fs, u) - f_qbar; sources
// outside solid x (or y) walls are not gas
// 16-bit storage: the projected links are stored rounded (dither of
// SALT_INIT), every other population keeps its stored bits
template <class S, class T>
__global__ __launch_bounds__(NT) void k_fs_project(Consts<T> c, const S* fin, S* fout, const T* ux, const T* uy, const T* uz,
                                                   const int* type, const int* obst, const T* fs_h) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t n = (size_t)c.nx * c.ny * c.nz;
    if (id >= n) return;
    int i, j, k;
    cell_ijk(c, id, i, j, k);
    for (int q = 0; q < 19; q++) fout[(size_t)q * n + id] = fin[(size_t)q * n + id];
    if (obst[id] != 0 || type[id] == LBM_FS_GAS) return;
    const size_t col = (size_t)j * c.nx + i;
    for (int q = 1; q < 19; q++) {
        int si = i - cxq(q), sj = j - cyq(q), sk = k - czq(q);
        bool fluid = true;
        if (c.x_boundary == LBM_XB_PERIODIC)
            si = pmod(si, c.nx);
        else if (si < 0 || si >= c.nx)
            fluid = false;
        if (c.y_boundary == LBM_YB_PERIODIC)
            sj = pmod(sj, c.ny);
        else if (sj < 0 || sj >= c.ny)
            fluid = false;
        bool gas = false;
        if (!fluid)
            gas = false;
        else if (sk >= c.nz)
            gas = true;
        else if (sk >= 0) {
            size_t sid = ((size_t)sk * c.ny + sj) * c.nx + si;
            gas = type[sid] == LBM_FS_GAS && obst[sid] == 0;
        }
        if (!gas) continue;
        const int qb = oppq(q);
        T rho_fs = c.rho0;
        if (c.pf != LBM_PF_TOTAL || c.fs_boundary_hydro) rho_fs = fs_density_link(c, k, q, true, type[id], fs_h, col);
        const T u = ux[id], v = uy[id], w = uz[id];
        const T fp = feq(rho_fs, u, v, w, q) + feq(rho_fs, u, v, w, qb) - Store<S, T>::load(fin, (size_t)qb * n + id, qb, c.store);
        if constexpr (std::is_same<S, T>::value)
            fout[(size_t)q * n + id] = fp;
        else
            Store<S, T>::store(fout, (size_t)q * n + id, q, fp, c.store, dither_base(SALT_INIT, (uint32_t)id));
    }
}

template <class S, class T>
static void init_t(const lbm_config_t& cfg, DevBufs& b, int slot, hipStream_t s) {
    Consts<T> c = make_consts<T>(cfg, -1);
    Slot<T> sl = make_slot<T>(b, slot);
    k_cellflags<T><<<nblocks(b.n), NT, 0, s>>>(c, sl, b.cellflags);
    k_init_populations<S, T><<<nblocks(b.n), NT, 0, s>>>(c, sl, (S*)b.f[0], (T*)b.rho, (T*)b.ux, (T*)b.uy, (T*)b.uz, b.obst,
                                                         b.type);
    if (cfg.fs_enabled && cfg.init_fs_populations) {
        k_fs_project<S, T><<<nblocks(b.n), NT, 0, s>>>(c, (const S*)b.f[0], (S*)b.f[1], (const T*)b.ux, (const T*)b.uy,
                                                        (const T*)b.uz, b.type, b.obst, (const T*)b.fs_h);
        HIP_OK(hipMemcpyAsync(b.f[0], b.f[1], 19 * b.n * sizeof(S), hipMemcpyDeviceToDevice, s));
    }
    if (cfg.fs_enabled && cfg.fs_tracking == LBM_FST_VOF) {
        VofArgs<T> a = make_vof<T>(cfg, b);
        k_vof_init_mass<T><<<nblocks(b.n), NT, 0, s>>>(a);
        k_vof_enforce<T><<<nblocks(b.n), NT, 0, s>>>(a);
        k_vof_height<T><<<(unsigned)((b.n2 + HC - 1) / HC), HT, 0, s>>>(c, sl, (const T*)b.fill, b.cellflags, (T*)b.fs_h);
    }
    HIP_OK(hipGetLastError());
}
void launch_init_populations(const lbm_config_t& c, DevBufs& b, int slot, hipStream_t s) {
    switch (b.storage) {
        case LBM_STORE_FP64: init_t<double, double>(c, b, slot, s); break;
        case LBM_STORE_FP16: init_t<__half, float>(c, b, slot, s); break;
        case LBM_STORE_BF16: init_t<bf16_t, float>(c, b, slot, s); break;
        default: init_t<float, float>(c, b, slot, s); break;
    }
}

// band: the passes skip the rows the narrow band shows settled (see
// band_mark); the marks of this step go to row-flag slot st.act_slot
template <class S, class T>
static void vof_pipeline(const lbm_config_t& cfg, DevBufs& b, S* f, int slot, int64_t step, bool band, StagedState& st,
                         hipStream_t s) {
    const size_t n = b.n;
    const unsigned nb = nblocks(n);
    const size_t rows = (size_t)cfg.nz * cfg.ny;
    int* act = b.rowact + (size_t)st.act_slot * rows;
    const int* act_prev = b.rowact + (size_t)(st.act_slot ^ 1) * rows;
    st.act_slot ^= 1;
    st.band_ok = 1;  // this step's marks will cover its end state
    auto A = [&]() {
        VofArgs<T> a = make_vof<T>(cfg, b);
        a.act = act;
        a.act_prev = act_prev;
        a.band = band ? 1 : 0;
        return a;
    };
    T empty_thr = (T)cfg.vof_empty_fill_threshold;
    k_vof_exchange<S, T><<<nb, NT, 0, s>>>(A(), f);
    k_vof_preclassify<T><<<nb, NT, 0, s>>>(A(), empty_thr);
    // a redistribution round and the classification that follows it; the
    // donor flags alternate between two slots (the preclassification and
    // every lone classification clear both)
    int rslot = 0;
    auto redistribute_finalize = [&]() {
        k_vof_redist_a<T><<<nb, NT, 0, s>>>(A(), rslot);
        k_vof_redist_fin<T><<<nb, NT, 0, s>>>(A(), empty_thr, rslot);
        std::swap(b.excess, b.excess2);
        rslot ^= 1;
    };
    for (int r = 0; r < 3; r++) redistribute_finalize();
    k_vof_enforce<T><<<nb, NT, 0, s>>>(A());
    T othr = (T)cfg.vof_orphan_empty_threshold;
    k_vof_orphan_a<T><<<nb, NT, 0, s>>>(A(), empty_thr, othr, cfg.vof_orphan_search_radius, cfg.vof_orphan_max_weak_neighbors);
    k_vof_orphan_b<T><<<nb, NT, 0, s>>>(A(), othr, cfg.vof_orphan_search_radius);
    k_vof_orphan_c<T><<<nb, NT, 0, s>>>(A());
    for (int mode = 0; mode < 2; mode++) {
        bool on = cfg.vof_detached_advection && (mode == 0 || cfg.vof_detached_residual_fill_threshold > 0.0f);
        if (on) {
            k_vof_detached_a<T><<<nb, NT, 0, s>>>(A(), mode, othr, cfg.vof_detached_max_wet_neighbors,
                                                  cfg.vof_detached_max_resolved_neighbors,
                                                  (T)cfg.vof_detached_residual_fill_threshold);
        }
        k_vof_detached_fin<T><<<nb, NT, 0, s>>>(A(), empty_thr, on ? mode : -1);
    }
    redistribute_finalize();
    k_vof_enforce<T><<<nb, NT, 0, s>>>(A());
    // collapse_trapped_gas_if_due (CODEX:2129-2143)
    if (cfg.vof_collapse_trapped_gas && step % cfg.vof_collapse_interval == 0) {
        k_flood_seed<T><<<nb, NT, 0, s>>>(A());
        // sweep flags: flagbuf[0] = 1 primes the first sweep; each sweep
        // clears its own flag before running
        int* flagbuf = b.si;  // reuse the int scratch: first (sweeps+1) ints
        // to the fixed point, bounded by 4 (nx + ny + nz) sweeps or by the
        // configured count if that is larger
        int sweeps = std::max(std::max(1, cfg.nx + cfg.ny + cfg.nz) * 4, cfg.vof_collapse_flood_sweeps);
        if ((size_t)(sweeps + 1) > n) sweeps = (int)n - 1;
        int one = 1;
        HIP_OK(hipMemcpyAsync(flagbuf, &one, sizeof(int), hipMemcpyHostToDevice, s));
        HIP_OK(hipMemsetAsync(flagbuf + 1, 0, sweeps * sizeof(int), s));
        // in batches of 16 sweeps; after each, the host reads the last
        // sweep's flag and stops at the fixed point (a sweep that finds
        // nothing to change returns at once, but thousands of launches
        // would still cost milliseconds).  Without flood_sync every sweep
        // is launched and the host never waits.
        for (int w = 0; w < sweeps;) {
            const int w1 = std::min(sweeps, w + 16);
            for (; w < w1; w++) k_flood_sweep<T><<<nb, NT, 0, s>>>(A(), flagbuf + w, flagbuf + w + 1);
            if (w < sweeps && st.flood_sync) {
                int changed = 1;
                HIP_OK(hipMemcpyAsync(&changed, flagbuf + w, sizeof(int), hipMemcpyDeviceToHost, s));
                HIP_OK(hipStreamSynchronize(s));
                if (!changed) break;
            }
        }
        HIP_OK(hipMemcpyAsync(b.icnt + IC_FLOOD_UNCONVERGED, flagbuf + sweeps, sizeof(int), hipMemcpyDeviceToDevice, s));
        const unsigned nch = (unsigned)reduce_chunks(n);
        k_collapse_sum<T><<<nch, RED_THREADS, 0, s>>>(A(), b.rpart);
        k_collapse_decide<<<1, RED_THREADS, 0, s>>>(b.rpart, (int)nch, b.dcnt, (double)cfg.vof_collapse_max_volume_m3);
        k_collapse_apply<T><<<nb, NT, 0, s>>>(A(), sweeps + 1);
        // (without a collapse the layer enforced just above holds: no LIQUID
        // cell touches a GAS one, so a second pass would change nothing)
        k_vof_enforce<T><<<nb, NT, 0, s>>>(A());
    }
    {
        Consts<T> c = make_consts<T>(cfg, -1);
        k_vof_height<T><<<(unsigned)((b.n2 + HC - 1) / HC), HT, 0, s>>>(c, make_slot<T>(b, slot), (const T*)b.fill, b.cellflags, (T*)b.fs_h);
    }
    k_vof_refill<S, T><<<nb, NT, 0, s>>>(A(), f, (uint32_t)step);
    k_vof_repair<S, T><<<nb, NT, 0, s>>>(A(), f);
    HIP_OK(hipGetLastError());
}

// S: the storage type (16-bit storage rounds at each of the step's stores:
// collision, stream, and the conditioning and refill where they write)
template <class S, class T>
static void staged_t(const lbm_config_t& cfg, DevBufs& b, int& cur, int slot, int64_t step, int omega_band_top, int count_links,
                     StagedState& st, hipStream_t s) {
    Consts<T> c = make_consts<T>(cfg, omega_band_top);
    Slot<T> sl = make_slot<T>(b, slot);
    const unsigned nb = nblocks(b.n);
    S* f = (S*)b.f[cur];
    S* fn = (S*)b.f[cur ^ 1];
    // a flat or piecewise bed (and the box) never moves: its cell flags and
    // obstacle masks are the same at every step once computed
    const bool static_geo = !c.bed_moving;
    const bool geo_cached = static_geo && st.geo_ready;
    // the conditioning rewrites populations in air runs and over a moving bed
    const bool cond_writes_f = c.air_walls || c.bed_moving;
    if (!geo_cached) k_cellflags<T><<<nb, NT, 0, s>>>(c, sl, b.cellflags);
    const bool height = cfg.fs_enabled && cfg.fs_tracking != LBM_FST_VOF;
    if (height) k_fs_classify<T><<<nb, NT, 0, s>>>(c, (const T*)b.fs_h, b.cellflags, (T*)b.fill, b.type, b.type_prev);
    // over a static geometry the obstacle masks settle after one pass
    // (obst_prev = obst = the closed mask), so without population writes
    // the conditioning has nothing left to do
    if (!geo_cached || cond_writes_f) k_cond<S, T><<<nb, NT, 0, s>>>(c, sl, f, b.obst, b.obst_prev, (uint32_t)step);
    st.geo_ready = static_geo;
    // a VOF run without conditioning: the stream of the last step left rho
    // and u as the macro step of its populations (see k_macro_fix)
    const bool vof_macro = cfg.fs_enabled && cfg.fs_tracking == LBM_FST_VOF && !cond_writes_f;
    if (vof_macro && st.macro_ready)
        k_macro_fix<S, T><<<nb, NT, 0, s>>>(c, f, (T*)b.rho, (T*)b.ux, (T*)b.uy, (T*)b.uz, b.type, b.type_prev, b.cellflags, b.mdirty,
                                            b.tstate);
    else
        k_macro<S, T><<<nb, NT, 0, s>>>(c, f, (T*)b.rho, (T*)b.ux, (T*)b.uy, (T*)b.uz, b.type, b.cellflags, c.fs_water ? b.tstate : nullptr);
    {
        // compiled per LES flag and collision model (fewer registers than
        // the kernel that holds every model)
        const int cm = collision_model(cfg);
#define LBM_COLLIDE(LES, CM)                                                                                                   \
    k_collide<S, T, LES, CM><<<nb, NT, 0, s>>>(c, sl, f, (const T*)b.rho, (const T*)b.ux, (const T*)b.uy, (const T*)b.uz, (T*)b.omega, \
                                                (T*)b.nu_t, (T*)b.smag, b.type, b.cellflags, (uint32_t)step)
        if (cfg.use_les) {
            if (cm == CM_REG) LBM_COLLIDE(true, CM_REG);
            else if (cm == CM_KBC) LBM_COLLIDE(true, CM_KBC);
            else LBM_COLLIDE(true, CM_ANY);
        } else {
            if (cm == CM_REG) LBM_COLLIDE(false, CM_REG);
            else if (cm == CM_KBC) LBM_COLLIDE(false, CM_KBC);
            else LBM_COLLIDE(false, CM_ANY);
        }
#undef LBM_COLLIDE
    }
    // the link counters of stream describe one step and are overwritten by
    // the next: they are counted on the last step of a call, the one a
    // caller can observe
    if (count_links) HIP_OK(hipMemsetAsync(b.icnt + IC_BOUNDARY_CUT, 0, 3 * sizeof(int), s));
    // a free-surface run takes the macro step of the streamed populations
    // inside the stream kernel
    auto stream = cfg.fs_enabled ? (count_links ? k_stream_bc<S, S, T, true, true> : k_stream_bc<S, S, T, false, true>)
                                 : (count_links ? k_stream_bc<S, S, T, true, false> : k_stream_bc<S, S, T, false, false>);
    // the stores the stream may leave out for long-standing gas cells (their
    // streak counts only where the macro fix runs; 16-bit stores differ from
    // step to step by their dither)
    const int gas_skip = (std::is_same<S, T>::value && static_geo && vof_macro && st.macro_ready) ? 1 : 0;
    stream<<<nb, NT, 0, s>>>(c, sl, f, fn, (const T*)b.ux, (const T*)b.uy, (const T*)b.uz, b.tstate, b.type, (const T*)b.fill,
                             (const T*)b.u_inlet, (const T*)b.inflow, (const T*)b.fs_h, b.cellflags, b.icnt, (T*)b.rho, (T*)b.ux,
                             (T*)b.uy, (T*)b.uz, (uint32_t)step, gas_skip);
    cur ^= 1;
    HIP_OK(hipGetLastError());
    if (cfg.fs_enabled) {
        S* f2 = (S*)b.f[cur];
        if (cfg.fs_tracking == LBM_FST_VOF) {
            // the band needs the marks of a whole previous pipeline and a
            // static geometry (cells that open or close are not marked)
            // (and a LIQUID cell, fill 1, counts as resolved in the orphan test)
            const bool band = st.band_allowed && st.band_ok && static_geo && cfg.vof_orphan_empty_threshold <= 1.0f;
            vof_pipeline<S, T>(cfg, b, f2, slot, step, band, st, s);
            st.macro_ready = vof_macro;
        } else {
            // height-function tracking (CODEX:3917-3921)
            const unsigned n2b = nblocks(b.n2);
            k_fs_flux<T><<<n2b, NT, 0, s>>>(c, (const T*)b.fill, b.type, b.cellflags, (const T*)b.ux, (const T*)b.uy, (T*)b.fs_qx,
                                            (T*)b.fs_qy);
            const T t = (T)(float)((double)(step + 1) * cfg.dt_phys);  // CODEX:3874
            k_fs_advect<T><<<n2b, NT, 0, s>>>(c, (const T*)b.fs_qx, (const T*)b.fs_qy, (const T*)b.h[slot], (T*)b.fs_h, t);
            k_fs_classify<T><<<nb, NT, 0, s>>>(c, (const T*)b.fs_h, b.cellflags, (T*)b.fill, b.type, b.type_prev);
            HIP_OK(hipMemsetAsync(b.icnt + IC_FS_REFILL, 0, sizeof(int), s));
            k_vof_refill<S, T><<<nb, NT, 0, s>>>(make_vof<T>(cfg, b), f2, (uint32_t)step);
            HIP_OK(hipGetLastError());
        }
    }
}
void launch_staged_step(const lbm_config_t& c, DevBufs& b, int& cur, int slot, int64_t step, int omega_band_top, int count_links,
                        StagedState& st, hipStream_t s) {
    switch (b.storage) {
        case LBM_STORE_FP64: staged_t<double, double>(c, b, cur, slot, step, omega_band_top, count_links, st, s); break;
        case LBM_STORE_FP16: staged_t<__half, float>(c, b, cur, slot, step, omega_band_top, count_links, st, s); break;
        case LBM_STORE_BF16: staged_t<bf16_t, float>(c, b, cur, slot, step, omega_band_top, count_links, st, s); break;
        default: staged_t<float, float>(c, b, cur, slot, step, omega_band_top, count_links, st, s); break;
    }
}

template <class S, class T>
static FusedArgs<S, T> make_fused_args(const lbm_config_t& cfg, DevBufs& b, int cur, int mode, int slot_prev, int slot,
                                       int64_t step, int write_macro, int omega_band_top) {
    FusedArgs<S, T> a;
    a.c = make_consts<T>(cfg, omega_band_top);
    a.s0 = make_slot<T>(b, slot_prev);
    a.s1 = make_slot<T>(b, slot);
    a.fin = (const S*)b.f[cur];
    a.fout = (S*)b.f[cur ^ 1];
    a.rho = (T*)b.rho;
    a.ux = (T*)b.ux;
    a.uy = (T*)b.uy;
    a.uz = (T*)b.uz;
    a.omega = (T*)b.omega;
    a.u_inlet = (const T*)b.u_inlet;
    a.inflow = (const T*)b.inflow;
    a.open = b.cellflags;
    a.mode = mode;
    a.write_macro = write_macro;
    a.step = (uint32_t)step;
    a.plane0 = 0;
    a.strips = 0;
    return a;
}

// the collision model is fixed at compile time unless the top sponge or
// backscatter can select another one per cell

// the tile flags of a case's wavefront step (gx x gy tiles per plane of
// bx x 256/bx cells)
static WaveQ make_waveq(const lbm_config_t& cfg, DevBufs& b, const FusedSched& fs, int bx, int gx, int gy) {
    WaveQ w;
    w.flag = b.waveq + WAVE_FLAG0;
    w.err = b.waveq;
    w.epoch = fs.epoch;
    w.bpp = gx * gy;
    const int lag = fs.lag > 0 ? fs.lag : (WAVE_OCC * fs.cus + w.bpp - 1) / w.bpp;
    w.lag = std::max(1, std::min(lag, cfg.nz));
    w.bx = bx;
    w.gx = gx;
    w.gy = gy;
    return w;
}

template <class S, class T>
static void fused_t(const lbm_config_t& cfg, DevBufs& b, int& cur, int mode, int slot_prev, int slot, int64_t step, int write_macro,
                    int omega_band_top, int band_top, int block_x, FusedSched& fs, hipStream_t s) {
    FusedArgs<S, T> a = make_fused_args<S, T>(cfg, b, cur, mode, slot_prev, slot, step, write_macro, omega_band_top);
    // planes 0 .. band_top can meet a cut or closed link (general kernels);
    // the planes above cannot (lean kernels)
    const int nb = band_top < 0 ? 0 : std::min(cfg.nz, band_top + 1);
    const int bxd = block_x > 0 ? block_x : 64;
    const dim3 blk(bxd, FB_THREADS / bxd);
    const unsigned gx = (cfg.nx + blk.x - 1) / blk.x, gy = (cfg.ny + blk.y - 1) / blk.y;
    const int cm = collision_model(cfg);
    a.band_planes = nb;
    // the XCD strips pay on the wavefront schedule (FP32: 6 % at 512x256x192,
    // its traffic 20 % lower); the VALU-bound 16-bit passes ran 1.5 % slower
    a.strips = fs.strips && gy % 8 == 0 && fs.wavefront && std::is_same<S, float>::value;
    // FP32 storage only: with FP64 the merged kernel spills to scratch, and
    // the VALU-bound 16-bit forms ran 20-30 % slower merged (SGPR spills)
    if constexpr (std::is_same<T, float>::value) {
    if (fs.wavefront && fs.resident) {
        // the tile-resident form: a co-resident grid of persistent blocks,
        // when it holds more blocks than a plane of tiles plus a row (or one
        // block per tile), and not inside a graph capture
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        HIP_OK(hipStreamIsCapturing(s, &cap));
        if (cap == hipStreamCaptureStatusNone) {
            // tiles of RES_THREADS cells: bx x (RES_THREADS / bx)
            const int rbx = std::min(bxd, RES_THREADS);
            const int rgx = (cfg.nx + rbx - 1) / rbx, rgy = (cfg.ny + RES_THREADS / rbx - 1) / (RES_THREADS / rbx);
            WaveQ w = make_waveq(cfg, b, fs, rbx, rgx, rgy);
            a.strips = fs.strips && rgy % 8 == 0;
            auto kern = cm == CM_REG ? k_step_resident<S, T, CM_REG>
                                     : (cm == CM_KBC ? k_step_resident<S, T, CM_KBC> : k_step_resident<S, T, CM_ANY>);
            int per_cu = 0;
            HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern), RES_THREADS, 0));
            int total = cfg.nz * w.bpp;
            const long long resident = (long long)per_cu * fs.cus;
            const int grid = (int)std::min<long long>(total, resident);
            // the longest forward wait, in tiles (a grid of one block per
            // tile never waits on a block that is not running)
            int reach = a.strips ? 2 * w.bpp : w.bpp + w.gx + 1;
            if (a.strips && !(grid == total || grid > reach)) {  // the plain order waits less far
                a.strips = 0;
                reach = w.bpp + w.gx + 1;
            }
            if (grid == total || grid > reach) {
                a.plane0 = 0;
                void* args[] = {&a, &w, &total};
                HIP_OK(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3(grid), dim3(RES_THREADS), args, 0, s));
                cur ^= 1;
                return;
            }
        }
    }
    }
    if (fs.wavefront && std::is_same<S, float>::value) {
        const WaveQ w = make_waveq(cfg, b, fs, bxd, (int)gx, (int)gy);
        auto kern = cm == CM_REG ? k_step_wave<S, T, CM_REG> : (cm == CM_KBC ? k_step_wave<S, T, CM_KBC> : k_step_wave<S, T, CM_ANY>);
        a.plane0 = 0;
        kern<<<2 * cfg.nz * w.bpp, FB_THREADS, 0, s>>>(a, w);
        HIP_OK(hipGetLastError());
        cur ^= 1;
        return;
    }
    if (nb > 0) {
        a.plane0 = 0;
        k_moments<S, T, true><<<dim3(gx, gy, nb), blk, 0, s>>>(a);
    }
    if (nb < cfg.nz) {
        a.plane0 = nb;
        k_moments<S, T, false><<<dim3(gx, gy, cfg.nz - nb), blk, 0, s>>>(a);
    }
    bool merged = false;
    // FP32 storage only: the FP16 / BF16 forms spill in the merged kernel
    if constexpr (std::is_same<S, float>::value && std::is_same<T, float>::value) {
        if (cm != CM_ANY && nb > 0 && nb < cfg.nz) {
            a.plane0 = 0;
            const dim3 grd(gx, gy, cfg.nz);
            if (cm == CM_REG)
                k_collide_all<S, T, CM_REG><<<grd, blk, 0, s>>>(a);
            else
                k_collide_all<S, T, CM_KBC><<<grd, blk, 0, s>>>(a);
            merged = true;
        }
    }
    if (!merged) {
        if (nb > 0) {
            a.plane0 = 0;
            const dim3 grd(gx, gy, nb);
            if (cm == CM_REG)
                k_collide<S, T, true, CM_REG><<<grd, blk, 0, s>>>(a);
            else if (cm == CM_KBC)
                k_collide<S, T, true, CM_KBC><<<grd, blk, 0, s>>>(a);
            else
                k_collide<S, T, true, CM_ANY><<<grd, blk, 0, s>>>(a);
        }
        if (nb < cfg.nz) {
            a.plane0 = nb;
            const dim3 grd(gx, gy, cfg.nz - nb);
            if (cm == CM_REG)
                k_collide<S, T, false, CM_REG><<<grd, blk, 0, s>>>(a);
            else if (cm == CM_KBC)
                k_collide<S, T, false, CM_KBC><<<grd, blk, 0, s>>>(a);
            else
                k_collide<S, T, false, CM_ANY><<<grd, blk, 0, s>>>(a);
        }
    }
    HIP_OK(hipGetLastError());
    cur ^= 1;
}
template <class S, class T>
static int fill_wave_case_t(const lbm_config_t& cfg, DevBufs& b, int cur, int mode, int slot_prev, int slot, int64_t step,
                            int write_macro, int omega_band_top, int band_top, int block_x, const FusedSched& fs, int block0, void* out) {
    WaveCase<S, T> e;
    memset(&e, 0, sizeof e);
    e.a = make_fused_args<S, T>(cfg, b, cur, mode, slot_prev, slot, step, write_macro, omega_band_top);
    e.a.band_planes = band_top < 0 ? 0 : std::min(cfg.nz, band_top + 1);
    const int bxd = block_x > 0 ? block_x : 64;
    const int gx = (cfg.nx + bxd - 1) / bxd, gy = (cfg.ny + FB_THREADS / bxd - 1) / (FB_THREADS / bxd);
    e.w = make_waveq(cfg, b, fs, bxd, gx, gy);
    e.w.lag = cfg.nz;  // the batch's order: every moments tile before every collision tile
    e.block0 = block0;
    memcpy(out, &e, sizeof e);
    return cfg.nz * e.w.bpp;  // blocks in each half of the grid
}
template <class S, class T>
static void patch_wave_case_t(void* entry, DevBufs& b, int cur, int mode, int slot_prev, int slot, int64_t step, int write_macro,
                              int epoch, int block0) {
    WaveCase<S, T>* e = (WaveCase<S, T>*)entry;
    e->a.s0 = make_slot<T>(b, slot_prev);
    e->a.s1 = make_slot<T>(b, slot);
    e->a.fin = (const S*)b.f[cur];
    e->a.fout = (S*)b.f[cur ^ 1];
    e->a.mode = mode;
    e->a.write_macro = write_macro;
    e->a.step = (uint32_t)step;
    e->w.epoch = epoch;
    e->block0 = block0;
}
void patch_wave_case(int storage, void* entry, DevBufs& b, int cur, int mode, int slot_prev, int slot, int64_t step, int write_macro,
                     int epoch, int block0) {
    switch (storage) {
        case LBM_STORE_FP64: patch_wave_case_t<double, double>(entry, b, cur, mode, slot_prev, slot, step, write_macro, epoch, block0); break;
        case LBM_STORE_FP16: patch_wave_case_t<__half, float>(entry, b, cur, mode, slot_prev, slot, step, write_macro, epoch, block0); break;
        case LBM_STORE_BF16: patch_wave_case_t<bf16_t, float>(entry, b, cur, mode, slot_prev, slot, step, write_macro, epoch, block0); break;
        default: patch_wave_case_t<float, float>(entry, b, cur, mode, slot_prev, slot, step, write_macro, epoch, block0); break;
    }
}

size_t wave_case_bytes(int storage) {
    switch (storage) {
        case LBM_STORE_FP64: return sizeof(WaveCase<double, double>);
        case LBM_STORE_FP16: return sizeof(WaveCase<__half, float>);
        case LBM_STORE_BF16: return sizeof(WaveCase<bf16_t, float>);
        default: return sizeof(WaveCase<float, float>);
    }
}
int wave_case_model(const lbm_config_t& c) { return collision_model(c); }
int fill_wave_case(const lbm_config_t& c, DevBufs& b, int cur, int mode, int slot_prev, int slot, int64_t step, int write_macro,
                   int omega_band_top, int band_top, int block_x, const FusedSched& fs, int block0, void* out) {
    if (!fused_supported(c)) throw std::runtime_error("fused path: needs air mode without a free surface");
    switch (b.storage) {
        case LBM_STORE_FP64: return fill_wave_case_t<double, double>(c, b, cur, mode, slot_prev, slot, step, write_macro, omega_band_top, band_top, block_x, fs, block0, out);
        case LBM_STORE_FP16: return fill_wave_case_t<__half, float>(c, b, cur, mode, slot_prev, slot, step, write_macro, omega_band_top, band_top, block_x, fs, block0, out);
        case LBM_STORE_BF16: return fill_wave_case_t<bf16_t, float>(c, b, cur, mode, slot_prev, slot, step, write_macro, omega_band_top, band_top, block_x, fs, block0, out);
        default: return fill_wave_case_t<float, float>(c, b, cur, mode, slot_prev, slot, step, write_macro, omega_band_top, band_top, block_x, fs, block0, out);
    }
}
template <class S, class T>
static void wave_batch_t(int cm, const void* tab, int ncases, int half, hipStream_t s) {
    auto kern = cm == CM_REG ? k_step_wave_batch<S, T, CM_REG> : (cm == CM_KBC ? k_step_wave_batch<S, T, CM_KBC> : k_step_wave_batch<S, T, CM_ANY>);
    kern<<<2 * half, FB_THREADS, 0, s>>>((const WaveCase<S, T>*)tab, ncases, half);
    HIP_OK(hipGetLastError());
}
void launch_wave_batch(int storage, int cm, const void* dev_table, int ncases, int half, hipStream_t s) {
    switch (storage) {
        case LBM_STORE_FP64: wave_batch_t<double, double>(cm, dev_table, ncases, half, s); break;
        case LBM_STORE_FP16: wave_batch_t<__half, float>(cm, dev_table, ncases, half, s); break;
        case LBM_STORE_BF16: wave_batch_t<bf16_t, float>(cm, dev_table, ncases, half, s); break;
        default: wave_batch_t<float, float>(cm, dev_table, ncases, half, s); break;
    }
}

void launch_fused_step(const lbm_config_t& c, DevBufs& b, int& cur, int mode, int slot_prev, int slot, int64_t step,
                       int write_macro, int omega_band_top, int band_top, int block_x, FusedSched& fs, hipStream_t s) {
    if (!fused_supported(c)) throw std::runtime_error("fused path: needs air mode without a free surface");
    switch (b.storage) {
        case LBM_STORE_FP64: fused_t<double, double>(c, b, cur, mode, slot_prev, slot, step, write_macro, omega_band_top, band_top, block_x, fs, s); break;
        case LBM_STORE_FP16: fused_t<__half, float>(c, b, cur, mode, slot_prev, slot, step, write_macro, omega_band_top, band_top, block_x, fs, s); break;
        case LBM_STORE_BF16: fused_t<bf16_t, float>(c, b, cur, mode, slot_prev, slot, step, write_macro, omega_band_top, band_top, block_x, fs, s); break;
        default: fused_t<float, float>(c, b, cur, mode, slot_prev, slot, step, write_macro, omega_band_top, band_top, block_x, fs, s); break;
    }
}



template <class S, class T>
static void materialize_t(const lbm_config_t& cfg, DevBufs& b, int cur, int slot, void* dst, hipStream_t s) {
    Consts<T> c = make_consts<T>(cfg, -1);
    Slot<T> sl = make_slot<T>(b, slot);
    k_cellflags<T><<<nblocks(b.n), NT, 0, s>>>(c, sl, b.cellflags);
    k_stream_bc<S, T, T, false, false><<<nblocks(b.n), NT, 0, s>>>(c, sl, (const S*)b.f[cur], (T*)dst, (const T*)b.ux, (const T*)b.uy,
                                                                   (const T*)b.uz, b.tstate, b.type, (const T*)b.fill, (const T*)b.u_inlet,
                                                                   (const T*)b.inflow, (const T*)b.fs_h, b.cellflags, b.icnt,
                                                                   nullptr, nullptr, nullptr, nullptr, 0u, 0);
    HIP_OK(hipGetLastError());
}
void launch_materialize(const lbm_config_t& c, DevBufs& b, int cur, int slot, void* dst, hipStream_t s) {
    switch (b.storage) {
        case LBM_STORE_FP64: materialize_t<double, double>(c, b, cur, slot, dst, s); break;
        case LBM_STORE_FP16: materialize_t<__half, float>(c, b, cur, slot, dst, s); break;
        case LBM_STORE_BF16: materialize_t<bf16_t, float>(c, b, cur, slot, dst, s); break;
        default: materialize_t<float, float>(c, b, cur, slot, dst, s); break;
    }
}

template <class S, class T>
static void store_pop_t(const lbm_config_t& cfg, DevBufs& b, int cur, const void* src, int64_t step, hipStream_t s) {
    Consts<T> c = make_consts<T>(cfg, -1);
    c.store_step = (uint32_t)step;
    k_store_populations<S, T><<<nblocks(19 * b.n), NT, 0, s>>>(c, (const T*)src, (S*)b.f[cur]);
    HIP_OK(hipGetLastError());
}
void launch_store_populations(const lbm_config_t& c, DevBufs& b, int cur, const void* src, int64_t step, hipStream_t s) {
    switch (b.storage) {
        case LBM_STORE_FP64: store_pop_t<double, double>(c, b, cur, src, step, s); break;
        case LBM_STORE_FP16: store_pop_t<__half, float>(c, b, cur, src, step, s); break;
        case LBM_STORE_BF16: store_pop_t<bf16_t, float>(c, b, cur, src, step, s); break;
        default: store_pop_t<float, float>(c, b, cur, src, step, s); break;
    }
}

size_t reduce_chunks(size_t n) { return (n + RED_CHUNK - 1) / RED_CHUNK; }

template <class S, class T>
static void reduce_t(const lbm_config_t& cfg, DevBufs& b, int cur, int src, int slot, int what, double* part, double* out,
                     hipStream_t s) {
    RedArgs<S, T> a;
    a.c = make_consts<T>(cfg, -1);
    a.s = make_slot<T>(b, slot);
    a.f = (const S*)b.f[cur];
    a.ux = (const T*)b.ux;
    a.uy = (const T*)b.uy;
    a.uz = (const T*)b.uz;
    a.mass = (const T*)b.mass;
    a.fill = (const T*)b.fill;
    a.u_inlet = (const T*)b.u_inlet;
    a.inflow = (const T*)b.inflow;
    a.fs_h = (const T*)b.fs_h;
    a.flags = b.cellflags;
    a.type = b.type;
    const unsigned nb = (unsigned)reduce_chunks(b.n);
    if (src)
        k_reduce_part<1, S, T><<<nb, RED_THREADS, 0, s>>>(a, what, part);
    else
        k_reduce_part<0, S, T><<<nb, RED_THREADS, 0, s>>>(a, what, part);
    k_reduce_final<<<1, RED_THREADS, 0, s>>>(part, (int)nb, 1, what == LBM_REDUCE_MAX_SPEED, out);
    HIP_OK(hipGetLastError());
}
void launch_reduce(const lbm_config_t& c, DevBufs& b, int cur, int post_stream, int slot, int what, double* part, size_t max_parts,
                   double* out, hipStream_t s) {
    if (reduce_chunks(b.n) > max_parts) throw std::runtime_error("reduction: more partial sums than the buffer holds");
    const int src = post_stream ? 0 : 1;
    switch (b.storage) {
        case LBM_STORE_FP64: reduce_t<double, double>(c, b, cur, src, slot, what, part, out, s); break;
        case LBM_STORE_FP16: reduce_t<__half, float>(c, b, cur, src, slot, what, part, out, s); break;
        case LBM_STORE_BF16: reduce_t<bf16_t, float>(c, b, cur, src, slot, what, part, out, s); break;
        default: reduce_t<float, float>(c, b, cur, src, slot, what, part, out, s); break;
    }
}

// The partial sums of lbm_reduce_noalloc live in this file-scope device
// array instead of a buffer of the handle: LBM_REDUCE_NOALLOC_MAX_CELLS /
// RED_CHUNK partials and one result slot.  Each device holds its own copy
// at its own address; host code reaches it only through the address that
// hipGetSymbolAddress returns for the current device (the array's name is a
// host-side stand-in, not a device address).
constexpr size_t NOALLOC_PARTS = (LBM_REDUCE_NOALLOC_MAX_CELLS + RED_CHUNK - 1) / RED_CHUNK;
static __device__ __align__(256) double scratch[NOALLOC_PARTS + 32];

double* noalloc_scratch(size_t* parts) {
    static std::mutex mu;
    static double* ptrs[256] = {};
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 256) throw std::runtime_error("device index out of range for the reduction scratch");
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        ptrs[dev] = (double*)p;
    }
    if (parts) *parts = NOALLOC_PARTS;
    return ptrs[dev];
}

template <class T>
static void cut_links_t(const lbm_config_t& cfg, DevBufs& b, int slot, hipStream_t s) {
    Consts<T> c = make_consts<T>(cfg, -1);
    HIP_OK(hipMemsetAsync(b.icnt + IC_BOUNDARY_CUT, 0, sizeof(int), s));
    k_cut_links<T><<<nblocks(b.n), NT, 0, s>>>(c, make_slot<T>(b, slot), b.icnt);
    HIP_OK(hipGetLastError());
}
void launch_cut_link_count(const lbm_config_t& c, DevBufs& b, int slot, hipStream_t s) {
    if (b.storage == LBM_STORE_FP64)
        cut_links_t<double>(c, b, slot, s);
    else
        cut_links_t<float>(c, b, slot, s);
}

template <class T>
__global__ __launch_bounds__(NT) void k_obstacle_from_flags(size_t n, const uint8_t* flags, int* obst) {
    size_t id = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < n) obst[id] = (flags[id] & 1u) ? 0 : 1;
}
void launch_obstacle(const lbm_config_t& c, DevBufs& b, int slot, int* dst, hipStream_t s) {
    launch_cellflags(c, b, slot, s);
    k_obstacle_from_flags<float><<<nblocks(b.n), NT, 0, s>>>(b.n, b.cellflags, dst);
    HIP_OK(hipGetLastError());
}

template <class S, class T>
static void load_pop_t(const lbm_config_t& cfg, DevBufs& b, int cur, void* dst, hipStream_t s) {
    Consts<T> c = make_consts<T>(cfg, -1);
    k_load_populations<S, T><<<nblocks(19 * b.n), NT, 0, s>>>(c, (const S*)b.f[cur], (T*)dst);
    HIP_OK(hipGetLastError());
}
void launch_load_populations(const lbm_config_t& c, DevBufs& b, int cur, void* dst, hipStream_t s) {
    switch (b.storage) {
        case LBM_STORE_FP64: load_pop_t<double, double>(c, b, cur, dst, s); break;
        case LBM_STORE_FP16: load_pop_t<__half, float>(c, b, cur, dst, s); break;
        case LBM_STORE_BF16: load_pop_t<bf16_t, float>(c, b, cur, dst, s); break;
        default: load_pop_t<float, float>(c, b, cur, dst, s); break;
    }
}

}  // namespace lbm
