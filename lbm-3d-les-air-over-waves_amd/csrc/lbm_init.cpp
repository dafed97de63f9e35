// lbm_init.cpp -- host half of the init sequence (CODEX:3777-3792): the
// macroscopic initial fields that lbm_initialize uploads before the device
// builds the populations (lbm_init_populations).  Host-only C++, no device
// calls: lbm_initial_fields can run on a machine without a GPU.
//
// Restated from the reference ("CODEX" = LBM_3D_LES_air_over_waves_CODEX.py):
//  * the log-law inlet profile and randomised velocity of init_fields
//    (CODEX:2714-2749);
//  * initialize_free_surface_height, classify_free_surface_cells and
//    apply_vof_initial_disconnected_shape (CODEX:1197-1287);
//  * prepare_initial_density_field (CODEX:743-818), hydrostatic and linear
//    wave;
//  * the water branch of init_fields (CODEX:2751-2800).
// The random draws follow NumPy's default_rng(seed) (SeedSequence + PCG64
// with XSL-RR output, uniform doubles from the top 53 bits, uniform floats
// from the top 24 bits of a 32-bit half), arrays in C order (nz, ny, nx).
// Arithmetic is float where the reference kernels are f32 and double where
// the host preparation is float64; transcendentals are the C library's double
// functions, rounded once to float where the value is f32, and sums run in
// index order, the same operations as the Python host init (lbm_init.py).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lbm.h"

namespace {

// ---- numpy.random.default_rng(seed): SeedSequence -> PCG64 ---------------
struct SeedSeq {
    uint32_t pool[4];
    static uint32_t hashmix(uint32_t v, uint32_t& hc) {
        v ^= hc;
        hc *= 0x931e8875u;
        v *= hc;
        v ^= v >> 16;
        return v;
    }
    static uint32_t mix(uint32_t x, uint32_t y) {
        uint32_t r = 0xca01f9ddu * x - 0x4973f715u * y;
        r ^= r >> 16;
        return r;
    }
    explicit SeedSeq(uint64_t seed) {
        std::vector<uint32_t> ent;
        ent.push_back((uint32_t)seed);
        if (seed >> 32) ent.push_back((uint32_t)(seed >> 32));
        uint32_t hc = 0x43b0d7e5u;
        for (int i = 0; i < 4; i++) pool[i] = hashmix(i < (int)ent.size() ? ent[i] : 0u, hc);
        for (int s = 0; s < 4; s++)
            for (int d = 0; d < 4; d++)
                if (s != d) pool[d] = mix(pool[d], hashmix(pool[s], hc));
        for (size_t s = 4; s < ent.size(); s++)
            for (int d = 0; d < 4; d++) pool[d] = mix(pool[d], hashmix(ent[s], hc));
    }
    void state64(uint64_t out[4]) const {
        uint32_t w[8];
        uint32_t hc = 0x8b51f9ddu;
        for (int i = 0; i < 8; i++) {
            uint32_t v = pool[i % 4];
            v ^= hc;
            hc *= 0x58f38dedu;
            v *= hc;
            v ^= v >> 16;
            w[i] = v;
        }
        for (int i = 0; i < 4; i++) out[i] = (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
    }
};

struct Pcg64 {
    unsigned __int128 state, inc;
    bool has32 = false;
    uint32_t u32 = 0;
    static unsigned __int128 mult() { return ((unsigned __int128)2549297995355413924ULL << 64) + 4865540595714422341ULL; }
    void step() { state = state * mult() + inc; }
    explicit Pcg64(uint64_t seed) {
        uint64_t v[4];
        SeedSeq(seed).state64(v);
        const unsigned __int128 s = ((unsigned __int128)v[0] << 64) + v[1];
        const unsigned __int128 q = ((unsigned __int128)v[2] << 64) + v[3];
        state = 0;
        inc = (q << 1) | 1u;
        step();
        state += s;
        step();
    }
    uint64_t next64() {
        step();
        const uint64_t x = (uint64_t)(state >> 64) ^ (uint64_t)state;
        const unsigned rot = (unsigned)(state >> 122);
        return (x >> rot) | (x << ((64u - rot) & 63u));
    }
    uint32_t next32() {
        if (has32) {
            has32 = false;
            return u32;
        }
        const uint64_t n = next64();
        has32 = true;
        u32 = (uint32_t)(n >> 32);
        return (uint32_t)(n & 0xffffffffu);
    }
    double next_double() { return (double)(next64() >> 11) * (1.0 / 9007199254740992.0); }
    float next_float() { return (float)(next32() >> 8) * (1.0f / 16777216.0f); }
    // Generator.uniform(low, high): low + (high - low) * U
    double uniform(double lo, double hi) { return lo + (hi - lo) * next_double(); }
};

// a float argument through the double C-library function, rounded once
float f32_exp(float x) { return (float)std::exp((double)x); }
float f32_log(float x) { return (float)std::log((double)x); }

struct Grid {
    int nx, ny, nz;
    size_t n2() const { return (size_t)nx * ny; }
    size_t n() const { return n2() * nz; }
};

// u_inlet_profile of init_fields (CODEX:2719-2738), f32
std::vector<float> inlet_profile(const lbm_init_params_t& p, int nz) {
    const float B = 5.0f;
    const float u_top = (float)p.u_top_lb;
    const float u_tau = 0.1f * u_top;
    const float nu = ((float)p.tau0 - 0.5f) / 3.0f;
    std::vector<float> tmp(nz), out(nz);
    for (int k = 0; k < nz; k++) {
        float yp = (float)k * u_tau / nu;
        if (yp <= 1.0f) yp = 1.0f;
        tmp[k] = (1.0f / (float)p.kappa) * f32_log(yp) + B;
    }
    const float top = tmp[nz - 1] != 0.0f ? tmp[nz - 1] : 1.0f;
    for (int k = 0; k < nz; k++) out[k] = tmp[k] / top * u_top;
    return out;
}

// log-law mean profile of the loglaw_noise state: A ln(max(z - z_b, z0) / z0)
// with the mean over the cells above the bed equal to inflow_u_lu (double)
std::vector<double> loglaw_profile(const lbm_init_params_t& p, int nz) {
    std::vector<double> prof(nz);
    const double z0 = p.inflow_z0_cells * p.dx;
    double sum = 0.0;
    int cnt = 0;
    for (int k = 0; k < nz; k++) {
        const double zr = ((double)k + 0.5) * p.dx - p.offset_m;
        prof[k] = std::log(std::max(zr, z0) / z0);
        if (zr > 0.0) {
            sum += prof[k];
            cnt++;
        }
    }
    const double mean = cnt ? sum / cnt : 1.0;
    for (int k = 0; k < nz; k++) prof[k] = p.inflow_u_lu * prof[k] / mean;
    return prof;
}

// x - x0 folded into [-Lx/2, Lx/2] on a periodic x axis, f32
float periodic_offset(const lbm_config_t& c, float xrel) {
    if (c.x_boundary != LBM_XB_PERIODIC) return xrel;
    const float lx = (float)c.nx * (float)(double)c.dx;
    if (xrel > 0.5f * lx) xrel = xrel - lx;
    if (xrel < -0.5f * lx) xrel = xrel + lx;
    return xrel;
}

// KdV solitary wave (CODEX:1166-1182) at t = 0: sech and tanh of its argument, f32
void solitary_profile(const lbm_config_t& c, const lbm_init_params_t& p, std::vector<float>& sech, std::vector<float>& tanh_) {
    const float dx = (float)p.dx;
    const float h0 = (float)p.free_surface_depth_m;
    const float amp = (float)p.solitary_amp_m;
    const float kappa = std::sqrt(3.0f * amp / (4.0f * h0 * h0 * h0));
    const float center = (float)p.solitary_x0_m + (float)p.solitary_direction * (float)p.solitary_celerity_phys * 0.0f;
    sech.resize(c.nx);
    tanh_.resize(c.nx);
    for (int i = 0; i < c.nx; i++) {
        const float xp = ((float)i + 0.5f) * dx;
        const float arg = kappa * periodic_offset(c, xp - center);
        const float ep = f32_exp(arg), em = f32_exp(-arg);
        const float den = ep + em;
        sech[i] = 2.0f / den;
        tanh_[i] = (ep - em) / den;
    }
}

// initialize_free_surface_height (CODEX:1197-1207), f32, (ny, nx)
std::vector<float> surface_height(const lbm_config_t& c, const lbm_init_params_t& p) {
    const Grid g{c.nx, c.ny, c.nz};
    std::vector<float> h(g.n2(), (float)p.free_surface_base_m), line(c.nx);
    const float dx = (float)p.dx;
    if (p.fs_initial == LBM_FSI_SOLITARY) {
        std::vector<float> sech, th;
        solitary_profile(c, p, sech, th);
        for (int i = 0; i < c.nx; i++) line[i] = (float)p.free_surface_base_m + (float)p.solitary_amp_m * sech[i] * sech[i];
    } else if (p.fs_initial == LBM_FSI_GAUSSIAN) {
        const float lx = (float)c.nx * dx;
        for (int i = 0; i < c.nx; i++) {
            const float xp = ((float)i + 0.5f) * dx;
            float xr = xp - (float)p.gaussian_center_x_m;
            if (c.x_boundary == LBM_XB_PERIODIC) {
                if (xr > 0.5f * lx) xr = xr - lx;
                if (xr < -0.5f * lx) xr = xr + lx;
            }
            const float arg = xr / (float)p.gaussian_sigma_m;
            line[i] = (float)p.free_surface_base_m + (float)p.gaussian_amp_m * f32_exp(-0.5f * arg * arg);
        }
    } else {
        return h;
    }
    for (int j = 0; j < c.ny; j++) memcpy(&h[(size_t)j * c.nx], line.data(), c.nx * sizeof(float));
    return h;
}

// classify_free_surface_cells + apply_vof_initial_disconnected_shape
// (CODEX:1209-1287), f32
void free_surface_state(const lbm_config_t& c, const lbm_init_params_t& p, const int32_t* obstacle0, std::vector<float>& fill,
                        std::vector<int32_t>& type, std::vector<float>& fs_h) {
    const Grid g{c.nx, c.ny, c.nz};
    const float dx = (float)p.dx;
    fs_h = surface_height(c, p);
    fill.assign(g.n(), 0.0f);
    type.assign(g.n(), LBM_FS_GAS);
    std::vector<float> fx, fy, fz;
    auto overlap = [&](int count, double bmin_d, double bmax_d, std::vector<float>& out) {
        const float bmin = (float)bmin_d, bmax = (float)bmax_d;
        out.resize(count);
        for (int a = 0; a < count; a++) {
            const float lo = (float)a * dx, hi = lo + dx;
            const float x0 = std::max(lo, bmin), x1 = std::min(hi, bmax);
            const float ov = std::max(x1 - x0, 0.0f);
            out[a] = ov / (hi - lo);
        }
    };
    if (p.vof_block) {
        overlap(c.nx, p.vof_block_center_m[0] - 0.5 * p.vof_block_size_m[0], p.vof_block_center_m[0] + 0.5 * p.vof_block_size_m[0], fx);
        overlap(c.ny, p.vof_block_center_m[1] - 0.5 * p.vof_block_size_m[1], p.vof_block_center_m[1] + 0.5 * p.vof_block_size_m[1], fy);
        overlap(c.nz, p.vof_block_center_m[2] - 0.5 * p.vof_block_size_m[2], p.vof_block_center_m[2] + 0.5 * p.vof_block_size_m[2], fz);
    }
    for (int k = 0; k < c.nz; k++)
        for (int j = 0; j < c.ny; j++)
            for (int i = 0; i < c.nx; i++) {
                const size_t col = (size_t)j * c.nx + i, id = (size_t)k * g.n2() + col;
                float f = (fs_h[col] - (float)k * dx) / dx;
                f = f < 0.0f ? 0.0f : (f > 1.0f ? 1.0f : f);
                int t = f <= 1e-4f ? LBM_FS_GAS : (f >= 1.0f - 1e-4f ? LBM_FS_LIQUID : LBM_FS_INTERFACE);
                const bool closed = obstacle0[id] != 0;
                if (closed) {
                    f = 0.0f;
                    t = LBM_FS_GAS;
                }
                if (p.vof_block && !closed) {
                    const float b = fx[i] * fy[j] * fz[k];
                    if (b > 1.0e-4f) {
                        float nf = std::max(f, b);
                        const bool liquid = nf >= 1.0f - 1.0e-4f;
                        if (liquid) nf = 1.0f;
                        f = nf;
                        t = liquid ? LBM_FS_LIQUID : LBM_FS_INTERFACE;
                    }
                }
                fill[id] = f;
                type[id] = t;
            }
}

// cosh(k zr) / cosh(k h) without overflow
double decay(double k, double zr, double depth) {
    if (!(k > 0.0)) return 1.0;
    const double kh = k * depth, kz = k * zr;
    return (std::exp(kz - kh) + std::exp(-kz - kh)) / (1.0 + std::exp(-2.0 * kh));
}

// pressure head of the linear potential-flow wave under eta(x) on a flat bed
// (CODEX:743-818, linear_wave): every horizontal mode of eta decays with the
// height above the bed as cosh(k z') / cosh(k h); cosine modes for solid x
// walls, periodic Fourier modes otherwise.  head[k][i], double.
std::vector<double> linear_wave_head(const lbm_config_t& c, const lbm_init_params_t& p, const std::vector<double>& eta, double bed) {
    const int nx = c.nx, nz = c.nz;
    const double dx = p.dx;
    const double depth = std::max(p.free_surface_base_m - bed, dx);
    const double PI = 3.14159265358979323846;
    std::vector<double> head((size_t)nz * nx);
    std::vector<double> x(nx);
    for (int i = 0; i < nx; i++) x[i] = ((double)i + 0.5) * dx;
    if (c.x_boundary == LBM_XB_SOLID) {
        std::vector<double> kn(nx), amp(nx, 0.0), basis((size_t)nx * nx);
        for (int m = 0; m < nx; m++) {
            kn[m] = (double)m * PI / (nx * dx);
            for (int i = 0; i < nx; i++) basis[(size_t)m * nx + i] = std::cos(kn[m] * x[i]);
        }
        double mean = 0.0;
        for (int i = 0; i < nx; i++) mean += eta[i];
        mean /= nx;
        for (int m = 0; m < nx; m++) {
            double s = 0.0;
            for (int i = 0; i < nx; i++) s += basis[(size_t)m * nx + i] * eta[i];
            amp[m] = (2.0 / nx) * s;
        }
        amp[0] = mean;
        for (int k = 0; k < nz; k++) {
            const double zc = ((double)k + 0.5) * dx;
            const double zr = std::min(std::max(zc - bed, 0.0), depth);
            for (int i = 0; i < nx; i++) {
                double s = 0.0;
                for (int m = 0; m < nx; m++) s += amp[m] * decay(kn[m], zr, depth) * basis[(size_t)m * nx + i];
                head[(size_t)k * nx + i] = s;
            }
        }
        return head;
    }
    // real DFT of eta (modes 0 .. nx/2), filtered per plane, inverse real DFT
    const int nm = nx / 2 + 1;
    std::vector<double> re(nm), im(nm), km(nm);
    for (int m = 0; m < nm; m++) {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < nx; i++) {
            const double ang = 2.0 * PI * (double)((long long)m * i % nx) / nx;
            a += eta[i] * std::cos(ang);
            b -= eta[i] * std::sin(ang);
        }
        re[m] = a;
        im[m] = b;
        km[m] = 2.0 * PI * (double)m / (nx * dx);
    }
    for (int k = 0; k < nz; k++) {
        const double zc = ((double)k + 0.5) * dx;
        const double zr = std::min(std::max(zc - bed, 0.0), depth);
        for (int i = 0; i < nx; i++) {
            double s = re[0] * decay(km[0], zr, depth);
            for (int m = 1; m < nm; m++) {
                const double d = decay(km[m], zr, depth);
                const double ang = 2.0 * PI * (double)((long long)m * i % nx) / nx;
                const bool nyq = (nx % 2 == 0) && m == nx / 2;
                const double term = d * (re[m] * std::cos(ang) - im[m] * std::sin(ang));
                s += nyq ? d * re[m] * std::cos(ang) : 2.0 * term;
            }
            head[(size_t)k * nx + i] = s / nx;
        }
    }
    return head;
}

// prepare_initial_density_field (CODEX:743-818), f32 result
std::vector<float> initial_density(const lbm_config_t& c, const lbm_init_params_t& p, const std::vector<float>& fs_h,
                                   const std::vector<double>& bed) {
    const Grid g{c.nx, c.ny, c.nz};
    std::vector<float> rho(g.n());
    bool flat = true;
    for (size_t a = 0; a < g.n2(); a++)
        if (std::fabs(bed[a] - bed[0]) > 1.0e-12) flat = false;
    if (p.linear_wave_pressure && p.fs_initial == LBM_FSI_GAUSSIAN && flat) {
        std::vector<double> eta(c.nx, 0.0);
        for (int i = 0; i < c.nx; i++) {
            double s = 0.0;
            for (int j = 0; j < c.ny; j++) s += (double)fs_h[(size_t)j * c.nx + i];
            eta[i] = s / c.ny - p.free_surface_base_m;
        }
        std::vector<double> head = linear_wave_head(c, p, eta, bed[0]);
        for (int k = 0; k < c.nz; k++)
            for (int i = 0; i < c.nx; i++) {
                double hd = head[(size_t)k * c.nx + i];
                if (p.hydrostatic_balanced != 1) {
                    const double z = ((double)k + 0.5) * p.dx;
                    hd = std::max(p.free_surface_base_m - z + hd, 0.0);
                }
                const float v = (float)(p.rho0 + p.g_lb * hd / p.dx / (1.0 / 3.0));
                for (int j = 0; j < c.ny; j++) rho[(size_t)k * g.n2() + (size_t)j * c.nx + i] = v;
            }
        return rho;
    }
    for (int k = 0; k < c.nz; k++) {
        const double zc = ((double)k + 0.5) * p.dx;
        for (size_t col = 0; col < g.n2(); col++) {
            const double h = (double)fs_h[col];
            const double head = (p.component_balanced == 1 || p.hydrostatic_balanced == 1) ? h - p.free_surface_base_m
                                                                                              : std::max(h - zc, 0.0);
            rho[(size_t)k * g.n2() + col] = (float)(p.rho0 + p.g_lb * head / p.dx / (1.0 / 3.0));
        }
    }
    return rho;
}

// initial velocity of the wet cells (CODEX:2756-2800), f32
void solitary_velocity(const lbm_config_t& c, const lbm_init_params_t& p, const std::vector<float>& fs_h, const std::vector<double>& bed,
                       std::vector<float>& ux, std::vector<float>& uz) {
    const Grid g{c.nx, c.ny, c.nz};
    ux.assign(g.n(), (float)p.u_current_lb);
    uz.assign(g.n(), 0.0f);
    if (p.fs_initial != LBM_FSI_SOLITARY || p.solitary_velocity == LBM_SOLVEL_NONE) return;
    const float h0 = (float)p.free_surface_depth_m;
    const float cc = (float)p.solitary_direction * (float)p.solitary_celerity_phys;
    const float vs = (float)p.vel_scale;
    std::vector<float> sech, th;
    if (p.solitary_velocity == LBM_SOLVEL_SHALLOW_DIVFREE) solitary_profile(c, p, sech, th);
    for (int j = 0; j < c.ny; j++)
        for (int i = 0; i < c.nx; i++) {
            const size_t col = (size_t)j * c.nx + i;
            float eta = fs_h[col] - (float)p.free_surface_base_m;
            if (eta < 0.0f) eta = 0.0f;
            const float u_sol = cc * eta / (h0 + eta);
            float du_dx = 0.0f;
            if (p.solitary_velocity == LBM_SOLVEL_SHALLOW_DIVFREE) {
                const float eta_x = -2.0f * (float)p.solitary_amp_m * (float)p.solitary_kappa_m_inv * sech[i] * sech[i] * th[i];
                du_dx = cc * h0 * eta_x / ((h0 + eta) * (h0 + eta));
            }
            for (int k = 0; k < c.nz; k++) {
                const size_t id = (size_t)k * g.n2() + col;
                ux[id] = ux[id] + u_sol / vs;
                if (p.solitary_velocity == LBM_SOLVEL_SHALLOW_DIVFREE) {
                    const float zc = ((float)k + 0.5f) * (float)p.dx;
                    float zr = zc - (float)bed[col];
                    if (zr < 0.0f) zr = 0.0f;
                    uz[id] = uz[id] + (-zr * du_dx) / vs;
                }
            }
        }
}

template <class R>
void put(void* dst, size_t i, double v) {
    ((R*)dst)[i] = (R)v;
}

}  // namespace

extern "C" int lbm_initial_fields(const lbm_config_t* c, const lbm_init_params_t* p, const int32_t* obstacle0, const void* bed,
                                  void* rho, void* ux, void* uy, void* uz, void* u_inlet, void* inflow, void* fill, int32_t* type,
                                  void* fs_h) {
    if (!c || !p || !obstacle0 || !rho || !ux || !uy || !uz || !u_inlet || !inflow) return LBM_E_ARG;
    if (p->version != LBM_INIT_PARAMS_VERSION || c->version != LBM_CONFIG_VERSION) return LBM_E_CONFIG;
    const bool water = c->fs_enabled != 0;
    if (water && (!fill || !type || !fs_h)) return LBM_E_ARG;
    const bool f64 = c->storage == LBM_STORE_FP64;
    auto set = [&](void* dst, size_t i, double v) { f64 ? put<double>(dst, i, v) : put<float>(dst, i, v); };
    auto get_bed = [&](size_t col) -> double {
        if (!bed) return (double)(float)p->offset_m;
        return f64 ? ((const double*)bed)[col] : (double)((const float*)bed)[col];
    };
    try {
        const Grid g{c->nx, c->ny, c->nz};
        const size_t n = g.n(), n2 = g.n2(), nz = (size_t)c->nz, ny = (size_t)c->ny;
        Pcg64 rng(p->seed);
        const std::vector<float> uin = inlet_profile(*p, c->nz);
        for (size_t k = 0; k < nz; k++) set(u_inlet, k, uin[k]);
        std::vector<double> vx(n), vy(n, 0.0), vz(n, 0.0), rh(n, p->rho0), infl(nz * ny, 0.0);
        if (p->initial_state == LBM_INIT_UNIFORM_NOISE) {
            const double a = p->init_noise_lu;
            for (size_t i = 0; i < n; i++) vx[i] = p->init_u_lu + rng.uniform(-a, a);
            for (size_t i = 0; i < n; i++) vy[i] = rng.uniform(-a, a);
            for (size_t i = 0; i < n; i++) vz[i] = rng.uniform(-a, a);
        } else if (p->initial_state == LBM_INIT_LOGLAW_NOISE) {
            const std::vector<double> prof = loglaw_profile(*p, c->nz);
            for (size_t i = 0; i < n; i++) vx[i] = prof[i / n2] * (1.0 + p->inflow_noise * rng.uniform(-1.0, 1.0));
            for (size_t i = 0; i < nz * ny; i++) infl[i] = prof[i / ny] * (1.0 + p->inflow_noise * rng.uniform(-1.0, 1.0));
        } else {
            // init_fields: ux = u_inlet_profile[k] * (1 + (2U - 1) TI), f32
            const float ti = (float)p->turbulence_intensity;
            for (size_t i = 0; i < n; i++) {
                const float kick = (rng.next_float() - 0.5f) * 2.0f;
                vx[i] = (double)(uin[i / n2] * (1.0f + kick * ti));
            }
        }
        if (water) {
            std::vector<float> fl, fsh, r0, u0, w0;
            std::vector<int32_t> ty;
            free_surface_state(*c, *p, obstacle0, fl, ty, fsh);
            std::vector<double> bd(n2);
            for (size_t a = 0; a < n2; a++) bd[a] = get_bed(a);
            r0 = initial_density(*c, *p, fsh, bd);
            solitary_velocity(*c, *p, fsh, bd, u0, w0);
            for (size_t i = 0; i < n; i++) {
                const bool gas = ty[i] == LBM_FS_GAS;
                rh[i] = gas ? p->rho0 : (double)r0[i];
                vx[i] = gas ? 0.0 : (double)u0[i];
                vy[i] = 0.0;
                vz[i] = gas ? 0.0 : (double)w0[i];
                set(fill, i, fl[i]);
                type[i] = ty[i];
            }
            for (size_t a = 0; a < n2; a++) set(fs_h, a, fsh[a]);
        }
        for (size_t i = 0; i < n; i++) {
            set(rho, i, rh[i]);
            set(ux, i, vx[i]);
            set(uy, i, vy[i]);
            set(uz, i, vz[i]);
        }
        for (size_t i = 0; i < nz * ny; i++) set(inflow, i, infl[i]);
    } catch (const std::exception&) {
        return LBM_E_NOMEM;
    }
    return LBM_OK;
}
