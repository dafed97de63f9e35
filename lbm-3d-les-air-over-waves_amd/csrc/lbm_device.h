// lbm_device.h -- device-side building blocks of the D3Q19 step (gfx950).
//
// Every function here restates one piece of the reference solver
// (LBM_3D_LES_air_over_waves_CODEX.py, "CODEX") with the same operation
// order, so that with -ffp-contract=off and IEEE division/sqrt the device
// results equal a serial CPU evaluation bit for bit.  T is the arithmetic
// type (float or double); S is the storage type of the populations.
#pragma once

#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/lbm.h"

namespace lbm {

// ---- D3Q19 tables (lbm_lattice.py:3-43) ---------------------------------
__host__ __device__ constexpr int cxq(int q) {
    return q == 1 || q == 7 || q == 10 || q == 11 || q == 14 ? 1 : (q == 2 || q == 8 || q == 9 || q == 12 || q == 13 ? -1 : 0);
}
__host__ __device__ constexpr int cyq(int q) {
    return q == 3 || q == 7 || q == 8 || q == 15 || q == 18 ? 1 : (q == 4 || q == 9 || q == 10 || q == 16 || q == 17 ? -1 : 0);
}
__host__ __device__ constexpr int czq(int q) {
    return q == 5 || q == 11 || q == 12 || q == 15 || q == 16 ? 1 : (q == 6 || q == 13 || q == 14 || q == 17 || q == 18 ? -1 : 0);
}
__host__ __device__ constexpr int oppq(int q) {
    return q == 0 ? 0 : (q <= 6 ? (q & 1 ? q + 1 : q - 1) : (q <= 10 ? (q <= 8 ? q + 2 : q - 2) : (q <= 14 ? (q <= 12 ? q + 2 : q - 2) : (q <= 16 ? q + 2 : q - 2))));
}
// z-mirror (free-slip top, new)
__host__ __device__ constexpr int mirzq(int q) {
    return q == 5 ? 6 : q == 6 ? 5 : q == 11 ? 14 : q == 14 ? 11 : q == 12 ? 13 : q == 13 ? 12 : q == 15 ? 18 : q == 18 ? 15 : q == 16 ? 17 : q == 17 ? 16 : q;
}
// index of the (cx,cy) column sample used by link q: (cx+1)*3 + (cy+1)
__host__ __device__ constexpr int c9q(int q) { return (cxq(q) + 1) * 3 + (cyq(q) + 1); }

template <class T>
__host__ __device__ constexpr T wq(int q) {
    return q == 0 ? T(1.0 / 3.0) : (q <= 6 ? T(1.0 / 18.0) : T(1.0 / 36.0));
}

// Division of a cell index u < 2^31 by an invariant d <= 2^31 as a
// multiply-high and a shift (Granlund and Montgomery, "Division by invariant
// integers using multiplication", 1994): with s = ceil(log2 d) and
// m = floor(2^(31+s) / d) + 1 < 2^32, floor(u / d) = (u m) >> (31 + s),
// because u m / 2^(31+s) exceeds u / d by less than 2^-s <= 1/d.
__host__ __device__ inline void fast_div_magic(uint32_t d, uint32_t& m, int& s) {
    s = 0;
    while (((uint64_t)1 << s) < d) s++;
    m = (uint32_t)(((uint64_t)1 << (31 + s)) / d + 1);
}
__host__ __device__ inline uint32_t fast_div(uint32_t u, uint32_t m, int s) {
    return (uint32_t)(((uint64_t)u * m) >> (31 + s));
}
// ---- storage types -------------------------------------------------------
struct bf16_t {
    uint16_t v;
};
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef __bf16 b2_t __attribute__((ext_vector_type(2)));

// Reduced-precision population codec (new).  A population is stored as
//   h_q = round((f_q - s_q) * 2^k),   f_q = s_q + h_q * 2^-k,
// with s_q = feq_q(1, u_ref) (s_q = w_q for u_ref = 0).  round() is either
// IEEE round-to-nearest-even or stochastic rounding driven by a counter-based
// dither of (cell, step, q): the 32-bit word r is added to the magnitude's
// bits below the kept precision before truncation, so a value is rounded up
// with probability equal to its distance from the lower neighbour.  The
// stochastic form removes the per-step rounding bias that otherwise
// accumulates into a momentum drift.
struct Codec {
    float scale, inv_scale;  // 2^k and 2^-k
    float shift[19];         // s_q
    int stochastic;          // LBM_ROUND_STOCHASTIC
};

__host__ __device__ inline uint32_t fmix32_hd(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
// per-cell dither base; population q uses base + q * 0x9E3779B9 (a Weyl
// sequence over q, so the 19 populations of a cell draw well-spread values)
__host__ __device__ inline uint32_t dither_base(uint32_t step, uint32_t cell) {
    return fmix32_hd(cell * 0x9E3779B1u ^ step * 0x85EBCA77u);
}
__host__ __device__ constexpr uint32_t dither_q(uint32_t base, int q) { return base + (uint32_t)q * 0x9E3779B9u; }
// the store points of a staged step draw their dither from the step number
// with a salt of their own (the collision's store, the only one of the fused
// step, keeps the plain step number)
constexpr uint32_t SALT_STREAM = 0x40000000u;  // post-stream populations
constexpr uint32_t SALT_REFILL = 0x20000000u;  // refilled free-surface cells
constexpr uint32_t SALT_COND = 0x10000000u;    // bed conditioning of closed links
constexpr uint32_t SALT_INIT = 0x08000000u;    // free-surface projection at initialisation

template <class S, class T>
struct Store;
template <class T>
struct Store<T, T> {  // full-precision storage: the population itself
    template <class I>
    static __device__ __forceinline__ T load(const T* p, I i, int, const Codec&) {
        return p[i];
    }
    // populations are stored once per step and read back only by the next
    // step: a non-temporal store keeps them from displacing the lines the
    // running step still reads from the L2
    template <class I>
    static __device__ __forceinline__ void store(T* p, I i, int, T v, const Codec&, uint32_t) {
        __builtin_nontemporal_store(v, p + i);
    }
    // store and return the value a later load reads back
    template <class I>
    static __device__ __forceinline__ T store_rt(T* p, I i, int, T v, const Codec&, uint32_t) {
        p[i] = v;
        return v;
    }
};
template <>
struct Store<__half, float> {  // (f_q - s_q) * 2^k as IEEE half
    // h * 2^-k is exact, so the fused multiply-add rounds exactly like
    // shift + h * 2^-k (one rounding, in the add); it issues as one mixed-
    // precision FMA that reads the half operand directly
    static __device__ __forceinline__ float decode_with(uint16_t bits, float shift, const Codec& k) {
        return __builtin_fmaf(__half2float(__ushort_as_half(bits)), k.inv_scale, shift);
    }
    static __device__ __forceinline__ float decode(uint16_t bits, int q, const Codec& k) { return decode_with(bits, k.shift[q], k); }
    static __device__ __forceinline__ uint16_t encode(float v, int q, const Codec& k, uint32_t base) {
        float d = (v - k.shift[q]) * k.scale;
        if (k.stochastic) {
            // the instruction rounds up iff r >> 19 plus the first 13 bits
            // below the half ulp carry into it (normal and subnormal range)
            h2_t old = {0, 0};
            h2_t hv = __builtin_amdgcn_cvt_sr_f16_f32(old, d, dither_q(base, q), false);
            return __builtin_bit_cast(uint16_t, hv[0]);
        }
        return __half_as_ushort(__float2half_rn(d));
    }
    template <class I>
    static __device__ __forceinline__ float load(const __half* p, I i, int q, const Codec& k) {
        return decode(__half_as_ushort(p[i]), q, k);
    }
    template <class I>
    static __device__ __forceinline__ void store(__half* p, I i, int q, float v, const Codec& k, uint32_t base) {
        __builtin_nontemporal_store(encode(v, q, k, base), reinterpret_cast<uint16_t*>(p) + i);  // as above
    }
    template <class I>
    static __device__ __forceinline__ float store_rt(__half* p, I i, int q, float v, const Codec& k, uint32_t base) {
        const uint16_t h = encode(v, q, k, base);
        reinterpret_cast<uint16_t*>(p)[i] = h;
        return decode(h, q, k);
    }
};
__device__ __forceinline__ float bf16_to_float(bf16_t b) { return __uint_as_float(((uint32_t)b.v) << 16); }
__device__ __forceinline__ bf16_t float_to_bf16(float f) {
    uint32_t u = __float_as_uint(f);
    bf16_t r;
    if ((u & 0x7fffffffu) > 0x7f800000u) {  // NaN stays NaN
        r.v = (uint16_t)((u >> 16) | 0x40u);
        return r;
    }
    u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
    r.v = (uint16_t)(u >> 16);
    return r;
}
// stochastic rounding to bfloat16: r >> 16 is added to the magnitude's bits
// below the bfloat16 ulp and the sum truncated (a carry may round up a
// binade; overflow saturates to infinity; NaN stays NaN, quieted).  This is
// exactly the rounding of v_cvt_sr_bf16_f32.
__device__ __forceinline__ bf16_t float_to_bf16_sr(float f, uint32_t r) {
    b2_t old = {0, 0};
    b2_t v = __builtin_amdgcn_cvt_sr_bf16_f32(old, f, r, false);
    bf16_t b;
    b.v = __builtin_bit_cast(uint16_t, v[0]);
    return b;
}
template <>
struct Store<bf16_t, float> {  // (f_q - s_q) * 2^k as bfloat16
    static __device__ __forceinline__ float decode_with(uint16_t bits, float shift, const Codec& k) {
        bf16_t b;
        b.v = bits;
        return shift + bf16_to_float(b) * k.inv_scale;
    }
    static __device__ __forceinline__ float decode(uint16_t bits, int q, const Codec& k) { return decode_with(bits, k.shift[q], k); }
    static __device__ __forceinline__ uint16_t encode(float v, int q, const Codec& k, uint32_t base) {
        float d = (v - k.shift[q]) * k.scale;
        return (k.stochastic ? float_to_bf16_sr(d, dither_q(base, q)) : float_to_bf16(d)).v;
    }
    template <class I>
    static __device__ __forceinline__ float load(const bf16_t* p, I i, int q, const Codec& k) {
        return decode(p[i].v, q, k);
    }
    template <class I>
    static __device__ __forceinline__ void store(bf16_t* p, I i, int q, float v, const Codec& k, uint32_t base) {
        __builtin_nontemporal_store(encode(v, q, k, base), reinterpret_cast<uint16_t*>(p) + i);  // as above
    }
    template <class I>
    static __device__ __forceinline__ float store_rt(bf16_t* p, I i, int q, float v, const Codec& k, uint32_t base) {
        const uint16_t h = encode(v, q, k, base);
        reinterpret_cast<uint16_t*>(p)[i] = h;
        return decode(h, q, k);
    }
};

// ---- deterministic elementary functions ---------------------------------
// Cody-Waite reduction by pi/2 + Taylor series (sin to r^17, cos to r^18),
// double precision, no contraction.  Same rule as the CPU oracle.
__device__ __forceinline__ double det_sincos(double x, int want_cos) {
    const double TWO_OVER_PI = 0.63661977236758134308;
    const double P1 = 1.57079632673412561417e+00;
    const double P2 = 6.07710050630396597660e-11;
    const double P3 = 2.02226624879595063154e-21;
    double n = floor(x * TWO_OVER_PI + 0.5);
    double r = x - n * P1;
    r = r - n * P2;
    r = r - n * P3;
    long long qi = (long long)n;
    int quad = (int)(((qi % 4) + 4) % 4);
    if (want_cos) quad = (quad + 1) & 3;
    double r2 = r * r;
    double s = 1.0 / 355687428096000.0;
    s = -1.0 / 1307674368000.0 + r2 * s;
    s = 1.0 / 6227020800.0 + r2 * s;
    s = -1.0 / 39916800.0 + r2 * s;
    s = 1.0 / 362880.0 + r2 * s;
    s = -1.0 / 5040.0 + r2 * s;
    s = 1.0 / 120.0 + r2 * s;
    s = -1.0 / 6.0 + r2 * s;
    double sinr = r + (r * r2) * s;
    double c = -1.0 / 6402373705728000.0;
    c = 1.0 / 20922789888000.0 + r2 * c;
    c = -1.0 / 87178291200.0 + r2 * c;
    c = 1.0 / 479001600.0 + r2 * c;
    c = -1.0 / 3628800.0 + r2 * c;
    c = 1.0 / 40320.0 + r2 * c;
    c = -1.0 / 720.0 + r2 * c;
    c = 1.0 / 24.0 + r2 * c;
    double cosr = (1.0 - 0.5 * r2) + (r2 * r2) * c;
    switch (quad) {
        case 0: return sinr;
        case 1: return cosr;
        case 2: return -sinr;
        default: return -cosr;
    }
}

// natural log: x = m 2^e, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s) as
// an odd series to s^23, plus e ln2 (two-part constant).  Same rule as the
// CPU oracle.
__device__ __forceinline__ double det_log(double x) {
    if (!(x > 0.0)) return x == 0.0 ? -__builtin_inf() : __builtin_nan("");
    if (__builtin_isinf(x)) return x;
    int e;
    double m = frexp(x, &e);
    if (m < 0.70710678118654752440) {
        m = m * 2.0;
        e = e - 1;
    }
    double s = (m - 1.0) / (m + 1.0);
    double s2 = s * s;
    double p = 1.0 / 23.0;
    p = 1.0 / 21.0 + s2 * p;
    p = 1.0 / 19.0 + s2 * p;
    p = 1.0 / 17.0 + s2 * p;
    p = 1.0 / 15.0 + s2 * p;
    p = 1.0 / 13.0 + s2 * p;
    p = 1.0 / 11.0 + s2 * p;
    p = 1.0 / 9.0 + s2 * p;
    p = 1.0 / 7.0 + s2 * p;
    p = 1.0 / 5.0 + s2 * p;
    p = 1.0 / 3.0 + s2 * p;
    double lm = 2.0 * s + (2.0 * s) * (s2 * p);
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    double de = (double)e;
    return (de * LN2_HI + lm) + de * LN2_LO;
}

/* exp(x) in double by Cody-Waite reduction by ln 2 (two-part constant) and a
   degree-14 Taylor polynomial on |r| <= ln2/2, scaled by 2^n; evaluated
   without contraction, so the device and the oracle agree bit for bit. */
__device__ __forceinline__ double det_exp(double x) {
    if (x != x) return x;
    if (x > 709.78) return __builtin_inf();
    if (x < -745.2) return 0.0;
    const double INV_LN2 = 1.44269504088896338700e+00;
    const double LN2_HI = 6.93147180369123816490e-01;
    const double LN2_LO = 1.90821492927058770002e-10;
    double n = floor(x * INV_LN2 + 0.5);
    double r = (x - n * LN2_HI) - n * LN2_LO;
    double p = 1.0 / 87178291200.0;    /* 1/14! */
    p = 1.0 / 6227020800.0 + r * p;    /* 1/13! */
    p = 1.0 / 479001600.0 + r * p;
    p = 1.0 / 39916800.0 + r * p;
    p = 1.0 / 3628800.0 + r * p;
    p = 1.0 / 362880.0 + r * p;
    p = 1.0 / 40320.0 + r * p;
    p = 1.0 / 5040.0 + r * p;
    p = 1.0 / 720.0 + r * p;
    p = 1.0 / 120.0 + r * p;
    p = 1.0 / 24.0 + r * p;
    p = 1.0 / 6.0 + r * p;
    p = 0.5 + r * p;
    p = 1.0 + r * p;
    p = 1.0 + r * p;
    return ldexp(p, (int)n);
}

__device__ __forceinline__ int pmod(int a, int n) {
    int r = a % n;
    return r < 0 ? r + n : r;
}
// pmod(a, n) for -n <= a < 2n (a neighbour one step away), without a division
__device__ __forceinline__ int wrap1(int a, int n) { return a < 0 ? a + n : (a >= n ? a - n : a); }

template <class T>
__device__ __forceinline__ T smoothstep01(T x) {
    return x * x * (T(3.0) - T(2.0) * x);
}

// calculate_feq (CODEX:623-643)
template <class T>
__device__ __forceinline__ T feq(T rho, T ux, T uy, T uz, int q) {
    T u2 = ux * ux + uy * uy + uz * uz;
    T cu = T(3.0) * ((T)cxq(q) * ux + (T)cyq(q) * uy + (T)czq(q) * uz);
    return rho * wq<T>(q) * (T(1.0) + cu + T(0.5) * cu * cu - T(1.5) * u2);
}

// periodic bilinear sample of a cell-centred (ny,nx) field (CODEX:2536-2569)
// bilinear sampling position of sample2d, split by axis: the x part depends
// on xq only and the y part on yq only, so the 9 link endpoints of a column
// need 3 + 3 of them, and several fields sampled at one point share them
// (same indices and weights, so the same results as sample2d)
template <class T>
struct Axis2d {
    int i0, i1;
    T f;
};
template <class T>
__device__ __forceinline__ Axis2d<T> sample_axis(T xq, int nx, T dx) {
    T Lx = (T)nx * dx;
    T Nxf = (T)nx;
    T xw = xq - floor(xq / Lx) * Lx;
    T u = xw / dx - T(0.5);
    u = u - floor(u / Nxf) * Nxf;
    int i0 = (int)floor(u);
    Axis2d<T> r;
    r.f = u - (T)i0;
    if (i0 >= nx) i0 -= nx;
    r.i0 = i0;
    r.i1 = i0 + 1 == nx ? 0 : i0 + 1;  // (i0 + 1) % nx for i0 in [-1, nx)
    return r;
}
template <class T>
__device__ __forceinline__ T sample_at(const T* H, const Axis2d<T>& x, const Axis2d<T>& y, int nx) {
    T h00 = H[(size_t)y.i0 * nx + x.i0], h10 = H[(size_t)y.i0 * nx + x.i1];
    T h01 = H[(size_t)y.i1 * nx + x.i0], h11 = H[(size_t)y.i1 * nx + x.i1];
    T hx0 = h00 * (T(1.0) - x.f) + h10 * x.f;
    T hx1 = h01 * (T(1.0) - x.f) + h11 * x.f;
    return hx0 * (T(1.0) - y.f) + hx1 * y.f;
}
template <class T>
__device__ __forceinline__ T sample2d(const T* H, T xq, T yq, int nx, int ny, T dx) {
    T Lx = (T)nx * dx, Ly = (T)ny * dx;
    T Nxf = (T)nx, Nyf = (T)ny;
    T xw = xq - floor(xq / Lx) * Lx;
    T yw = yq - floor(yq / Ly) * Ly;
    T u = xw / dx - T(0.5);
    T v = yw / dx - T(0.5);
    u = u - floor(u / Nxf) * Nxf;
    v = v - floor(v / Nyf) * Nyf;
    int i0 = (int)floor(u), j0 = (int)floor(v);
    T fx = u - (T)i0, fy = v - (T)j0;
    if (i0 >= nx) i0 -= nx;
    if (j0 >= ny) j0 -= ny;
    int i1 = (i0 + 1) % nx, j1 = (j0 + 1) % ny;
    T h00 = H[(size_t)j0 * nx + i0], h10 = H[(size_t)j0 * nx + i1];
    T h01 = H[(size_t)j1 * nx + i0], h11 = H[(size_t)j1 * nx + i1];
    T hx0 = h00 * (T(1.0) - fx) + h10 * fx;
    T hx1 = h01 * (T(1.0) - fx) + h11 * fx;
    return hx0 * (T(1.0) - fy) + hx1 * fy;
}

// ---- per-run constants ---------------------------------------------------
template <class T>
struct Consts {
    int nx, ny, nz;
    uint32_t div_nx_m, div_ny_m;  // cell index -> (i, j, k) by multiply-high (see cell_ijk)
    int div_nx_s, div_ny_s;
    int fs_water;    // a free surface is tracked (the reference's water_free_surface_enabled)
    int air_walls;   // air physics: closed links of static walls are conditioned too (CODEX:2922)
    int collision_reg, use_les;
    int x_boundary, y_boundary, top_boundary;
    int bed_profile, bed_moving, include_bed, wall_link;
    int top_drive, x_fringe, top_sponge;
    int fs_vof, pf, fs_boundary_hydro, refill_mode, link_aperture, thin_gap_bridge;
    int fs_mode;  // LBM_FSM_*
    int omega_band_top;  // highest k whose omega must persist (moving beds), -1 otherwise
    uint32_t seed;
    T dx, rho0, tau0, nu0, cs2_les, nu_t_max, kappa, z0_phys;
    T u_top_lb, u_current_lb, g_lb, g_body_lb, c_backscatter;
    T free_surface_base_m, bridge_strength;
    T offset_m, bed_amp_m, bed_wavelength_m, wave_speed, wave_period, depth_swl_m, vel_scale, phase_speed;
    int include_cube;
    int phi_fraction, phi_flags;  // open fractions / open flags from phi sampled at the link endpoints
    T cube_half[3], cube_center[3], cube_rot[4];  // box obstacle: half sizes, centre, cos/sin yaw, cos/sin pitch
    T solitary[5];                                // amplitude, depth, x0, direction, celerity
    T dt_phys;
    T inv_dx;     // 1/dx, used in place of a division by dx when dx is a power of two (exact)
    int dx_pow2;
    Codec store;          // reduced-precision population codec
    uint32_t store_step;  // step index of the dither for stores outside the step kernels
};

// x / dx, exactly: a product with 1/dx when dx is a power of two
template <class T>
__device__ __forceinline__ T div_dx(const Consts<T>& c, T x) {
    return c.dx_pow2 ? x * c.inv_dx : x / c.dx;
}

// 2-D time-slot arrays produced by the bed kernels for one time level
template <class T>
struct Slot {
    const T* h;      // water_h (bed height)
    const T* uw;     // u_wave
    const T* vw;     // v_wave
    const T* ww;     // w_wave
    const T* eta9;   // [9][ny][nx] bed height at the 9 link endpoints of each column
    const T* denom;  // [ny][nx] slope normaliser of phi
    const int* kband;   // [ny][nx] highest k with a cut or closed link
    const T* uw9;       // [3][9][ny][nx] wall velocity at the 9 link endpoints
    const int* ksolid;  // [ny][nx] highest k whose 18 links the bed closes (-1: none)
    const T* fsol;      // [19][ny][nx] feq(rho0, wall velocity of link q, q): the conditioned
                        // value of a closed link, for every column
    const T* ibb;       // [18][ny][nx] moving-wall term 6 w_qb rho0 (c_qb . u_wall) of link q = 1..18
    const T* phi;       // [nz][ny][nx] phi of this time level (phi boundary-geometry modes only)
};

// link geometry of one cell (CODEX:2492-2686): open flag, open fraction,
// wall type for every q, plus the cell flags lo[0] and near_obstacle.
template <class T>
struct CellGeo {
    uint32_t open_mask;  // bit q set when lattice_open[q] == 1 (bit 0 = any open)
    uint32_t xwall_mask; // bit q set when the link is an x/y wall override
    uint32_t cube_mask;  // bit q set when the box obstacle is the nearer solid of link q (bit 0: cell centre)
    int near;
    T lf[19];
};

// signed distance to the box obstacle (CODEX:897-934): offsets folded into
// the periodic domain (x only when x is periodic, y always), rotated by the
// yaw about z and the pitch about y, then the box distance (outside: length
// of the positive excess; inside: the largest, negative, excess)
template <class T>
__device__ __forceinline__ T cube_sdf(const Consts<T>& c, T xp, T yp, T zp) {
    const T lx = (T)c.nx * c.dx, ly = (T)c.ny * c.dx;
    T dxp = xp - c.cube_center[0];
    T dyp = yp - c.cube_center[1];
    if (c.x_boundary == LBM_XB_PERIODIC) {
        if (dxp > T(0.5) * lx) dxp -= lx;
        if (dxp < T(-0.5) * lx) dxp += lx;
    }
    if (dyp > T(0.5) * ly) dyp -= ly;
    if (dyp < T(-0.5) * ly) dyp += ly;
    const T yc = c.cube_rot[0], ys = c.cube_rot[1], pc = c.cube_rot[2], ps = c.cube_rot[3];
    T xl = yc * dxp + ys * dyp;
    T yl = -ys * dxp + yc * dyp;
    T z1 = zp - c.cube_center[2];
    T xpitch = pc * xl - ps * z1;
    T zl = ps * xl + pc * z1;
    T qx = fabs(xpitch) - c.cube_half[0];
    T qy = fabs(yl) - c.cube_half[1];
    T qz = fabs(zl) - c.cube_half[2];
    T ox = qx > T(0) ? qx : T(0), oy = qy > T(0) ? qy : T(0), oz = qz > T(0) ? qz : T(0);
    T outside = sqrt(ox * ox + oy * oy + oz * oz);
    T m = qy > qz ? qy : qz;
    m = qx > m ? qx : m;
    T inside = m < T(0) ? m : T(0);
    return outside + inside;
}

// sample_phi_periodic (CODEX:661-707): trilinear sample of phi, periodic
// bilinear in x / y with the bed sampling's indices, clamped in z
template <class T>
__device__ __forceinline__ T sample_phi(const Consts<T>& c, const T* phi, T xp, T yp, T zp) {
    const T dx = c.dx;
    const Axis2d<T> ax = sample_axis(xp, c.nx, dx), ay = sample_axis(yp, c.ny, dx);
    T r = zp / dx - T(0.5);
    int k0 = (int)floor(r);
    T fz = r - (T)k0;
    int k1 = k0 + 1;
    if (k0 < 0) {
        k0 = 0;
        k1 = 0;
        fz = T(0);
    }
    if (k1 >= c.nz) {
        k0 = c.nz - 1;
        k1 = c.nz - 1;
        fz = T(0);
    }
    const size_t plane = (size_t)c.nx * c.ny;
    const T p0 = sample_at(phi + (size_t)k0 * plane, ax, ay, c.nx);
    const T p1 = sample_at(phi + (size_t)k1 * plane, ax, ay, c.nx);
    return p0 * (T(1.0) - fz) + p1 * fz;
}

// open fraction of the cell centre (link 0) and whether the box is the
// nearer surface there
template <class T>
__device__ __forceinline__ T cell_center_fraction(const Consts<T>& c, const Slot<T>& s, int i, int j, int k, bool& by_cube) {
    const size_t col = (size_t)j * c.nx + i;
    const size_t plane = (size_t)c.nx * c.ny;
    const T dx = c.dx;
    const T zc = ((T)k + T(0.5)) * dx;
    const T xc = ((T)i + T(0.5)) * dx, yc = ((T)j + T(0.5)) * dx;
    by_cube = false;
    T eta_c = s.eta9[4 * plane + col];
    T gap_c = zc - eta_c;
    T ggap = c.include_bed ? gap_c : T(1.0e6);
    if (c.include_cube) {
        T cg = cube_sdf(c, xc, yc, zc);
        if (cg < ggap) {
            ggap = cg;
            by_cube = true;
        }
    }
    T sgap = c.phi_fraction ? s.phi[((size_t)k * c.ny + j) * c.nx + i] : ggap;
    T frac = T(0.5) + div_dx(c, sgap);
    if (frac < T(0)) frac = T(0);
    if (frac > T(1)) frac = T(1);
    return frac;
}

template <class T>
// exact_walls: also the box bit of every open link (the wall type of each
// link, for the geometry download); the step kernels read the wall type of
// closed links only
__device__ __forceinline__ void cell_geometry(const Consts<T>& c, const Slot<T>& s, int i, int j, int k, CellGeo<T>& g,
                                              bool exact_walls = false) {
    const size_t col = (size_t)j * c.nx + i;
    const size_t plane = (size_t)c.nx * c.ny;
    const T dx = c.dx;
    T zc = ((T)k + T(0.5)) * dx;
    const T xc = ((T)i + T(0.5)) * dx, yc = ((T)j + T(0.5)) * dx;
    uint32_t mask = 0, xw = 0, cube = 0;
    // a cell at least 2 dx above the highest of its 9 bed samples and at
    // least 2 dx outside the box: every link endpoint lies within 0.71 dx of
    // the centre and both distances change by at most that much (the box
    // distance is a Euclidean distance), so every gap is >= dx/2, every
    // fraction clamps to 1 and every link is open; the solid x / y wall
    // overrides below then close the links leaving the domain (the values of
    // the per-link evaluation, without its bed and box work)
    if (!c.phi_fraction && !(c.include_cube && exact_walls)) {
        T e9[9];
#pragma unroll
        for (int e = 0; e < 9; e++) e9[e] = s.eta9[(size_t)e * plane + col];  // issued together
        bool far = true;  // false for a NaN sample, which takes the per-link path
#pragma unroll
        for (int e = 0; e < 9; e++) far = far & (zc - e9[e] >= T(2.0) * dx);
        const bool far_box = !c.include_cube || cube_sdf(c, xc, yc, zc) >= T(2.0) * dx;
        if ((!c.include_bed || far) && far_box) {
            g.lf[0] = T(1);
#pragma unroll
            for (int q = 1; q < 19; q++) {
                bool wall = false;
                if (c.x_boundary == LBM_XB_SOLID) wall |= (i == 0 && cxq(q) > 0) || (i == c.nx - 1 && cxq(q) < 0);
                if (c.y_boundary == LBM_YB_SOLID) wall |= (j == 0 && cyq(q) > 0) || (j == c.ny - 1 && cyq(q) < 0);
                g.lf[q] = wall ? T(0) : T(1);
                xw |= (uint32_t)wall << q;
            }
            mask = 0x7fffeu & ~xw;
            // a cell has at most 9 wall links (corner column), so one stays open
            g.near = mask != 0x7fffeu;
            g.open_mask = mask | 1u;
            g.xwall_mask = xw;
            g.cube_mask = 0;
            return;
        }
    }
    {
        bool by_cube;
        g.lf[0] = cell_center_fraction(c, s, i, j, k, by_cube);
        if (by_cube) cube |= 1u;
    }
#pragma unroll
    for (int q = 1; q < 19; q++) {
        T zq = zc + T(0.5) * dx * (T)czq(q);
        T eta = s.eta9[(size_t)c9q(q) * plane + col];
        T gap = zq - eta;
        T ggap = c.include_bed ? gap : T(1.0e6);
        if (c.include_cube) {
            T cg = cube_sdf(c, xc + T(0.5) * dx * (T)cxq(q), yc + T(0.5) * dx * (T)cyq(q), zq);
            if (cg < ggap) {
                ggap = cg;
                cube |= 1u << q;
            }
        }
        T sgap = c.phi_fraction ? sample_phi(c, s.phi, xc + T(0.5) * dx * (T)cxq(q), yc + T(0.5) * dx * (T)cyq(q), zq) : ggap;
        T frac = T(0.5) + div_dx(c, sgap);
        if (frac < T(0)) frac = T(0);
        if (frac > T(1)) frac = T(1);
        int lo = (c.phi_flags ? sgap : ggap) < T(0) ? 0 : 1;
        if (c.x_boundary == LBM_XB_SOLID) {
            if ((i == 0 && cxq(q) > 0) || (i == c.nx - 1 && cxq(q) < 0)) {
                lo = 0;
                frac = T(0);
                xw |= 1u << q;
            }
        }
        if (c.y_boundary == LBM_YB_SOLID) {
            if ((j == 0 && cyq(q) > 0) || (j == c.ny - 1 && cyq(q) < 0)) {
                lo = 0;
                frac = T(0);
                xw |= 1u << q;
            }
        }
        g.lf[q] = frac;
        mask |= (uint32_t)lo << q;
    }
    int any_open = (mask & 0x7fffeu) != 0;
    g.near = (mask & 0x7fffeu) != 0x7fffeu;
    if (!any_open) {
#pragma unroll
        for (int q = 1; q < 19; q++) g.lf[q] = T(0);
    }
    g.open_mask = mask | (uint32_t)any_open;
    g.xwall_mask = xw;
    g.cube_mask = cube;
}

// wall type of link q given the geometry (bed-only obstacle): x/y wall
// overrides first, otherwise the bed when included
template <class T>
__device__ __forceinline__ int link_wall_type(const Consts<T>& c, const CellGeo<T>& g, int i, int j, int q) {
    if (q > 0 && ((g.xwall_mask >> q) & 1u)) {
        // the y override is applied after the x override, so it wins
        if (c.y_boundary == LBM_YB_SOLID && j == 0 && cyq(q) > 0) return LBM_WALL_YMIN;
        if (c.y_boundary == LBM_YB_SOLID && j == c.ny - 1 && cyq(q) < 0) return LBM_WALL_YMAX;
        return (i == 0 && cxq(q) > 0) ? LBM_WALL_XMIN : LBM_WALL_XMAX;
    }
    if ((g.cube_mask >> q) & 1u) return LBM_WALL_CUBE;
    return c.include_bed ? LBM_WALL_BED : LBM_WALL_NONE;
}

// wall velocity of link q (CODEX:937-988 link sampling, or the cell value)
template <class T>
__device__ __forceinline__ void wall_velocity(const Consts<T>& c, const Slot<T>& s, int i, int j, int q, T& ux, T& uy, T& uz) {
    if (c.wall_link) {
        const size_t plane2 = (size_t)c.nx * c.ny;
        const size_t col = (size_t)j * c.nx + i;
        const int e = c9q(q);
        ux = s.uw9[(size_t)e * plane2 + col];
        uy = s.uw9[(size_t)(9 + e) * plane2 + col];
        uz = s.uw9[(size_t)(18 + e) * plane2 + col];
    } else {
        size_t col = (size_t)j * c.nx + i;
        ux = s.uw[col];
        uy = s.vw[col];
        uz = s.ww[col];
    }
}

// phi at cell centre (CODEX:2432-2490): per-column slope normaliser + x wall
template <class T>
__device__ __forceinline__ T cell_phi(const Consts<T>& c, const Slot<T>& s, int i, int j, int k) {
    size_t col = (size_t)j * c.nx + i;
    T dx = c.dx;
    T zc = ((T)k + T(0.5)) * dx;
    T xc = ((T)i + T(0.5)) * dx;
    T phi = T(1.0e6);
    if (c.include_bed) phi = (zc - s.h[col]) / s.denom[col];
    if (c.include_cube) {
        T cp = cube_sdf(c, xc, ((T)j + T(0.5)) * dx, zc);
        if (cp < phi) phi = cp;
    }
    if (c.x_boundary == LBM_XB_SOLID) {
        T a = xc, b = (T)c.nx * dx - xc;
        T xw = a < b ? a : b;
        if (xw < phi) phi = xw;
    }
    return phi;
}

// update_bed_populations_and_reinit for one cell (CODEX:2894-2940)
template <class T>
__device__ __forceinline__ void condition_cell(const Consts<T>& c, const Slot<T>& s, const CellGeo<T>& g, int i, int j, T f[19]) {
    const bool air = c.air_walls;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        bool open = q == 0 ? (g.open_mask & 1u) : ((g.open_mask >> q) & 1u);
        if (open) continue;
        int wt = link_wall_type(c, g, i, j, q);
        T ux = T(0), uy = T(0), uz = T(0);
        if (wt == LBM_WALL_BED && c.bed_moving) {
            wall_velocity(c, s, i, j, q, ux, uy, uz);
        } else if (air) {
            if (wt == LBM_WALL_BED) wall_velocity(c, s, i, j, q, ux, uy, uz);
        } else {
            continue;
        }
        T fe = feq(c.rho0, ux, uy, uz, q);
        T frac = g.lf[q];
        f[q] = frac * f[q] + (T(1.0) - frac) * fe;
    }
}

// the moving-wall IBB of apply_open_boundary_conditions (CODEX:3674-3712)
// for a near-obstacle cell; fpre = the cell's post-collision populations
template <class T>
__device__ __forceinline__ T ibb_link(const Consts<T>& c, const Slot<T>& s, const CellGeo<T>& g, int i, int j, int q, T fnew, T fpre_qb) {
    const int qb = oppq(q);
    T fo = g.lf[q];
    T ux = T(0), uy = T(0), uz = T(0);
    if (link_wall_type(c, g, i, j, q) == LBM_WALL_BED) wall_velocity(c, s, i, j, q, ux, uy, uz);
    T cuw = (T)cxq(qb) * ux + (T)cyq(qb) * uy + (T)czq(qb) * uz;
    T fw = fpre_qb - T(6.0) * wq<T>(qb) * c.rho0 * cuw;
    return fnew * fo + fw * (T(1.0) - fo);
}

// counter-based uniform [0,1) for backscatter (same hash as the oracle)
__device__ __forceinline__ float bs_uniform(uint32_t seed, uint32_t step, uint32_t cell) {
    uint32_t h = fmix32_hd(seed ^ fmix32_hd(cell + fmix32_hd(step * 0x9E3779B9u)));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

// compute_LES strain -> omega for one active cell (CODEX:3040-3141).  The
// neighbour velocities are passed in; wall model inputs optional.
template <class T>
__device__ __forceinline__ T les_omega(const Consts<T>& c, T uxp, T uxm, T uyp_, T uym_, T uzp_, T uzm_,  // ux at x+,x-,y+,y-,z+,z-
                                       T vxp, T vxm, T vyp, T vym, T vzp, T vzm,                        // uy
                                       T wxp, T wxm, T wyp, T wym, T wzp, T wzm,                        // uz
                                       bool near_wall, T phi, T u_here, T v_here, T uwall_x, T uwall_y, T* nu_t_out, T* smag_out) {
    T du_dx = T(0.5) * (uxp - uxm);
    T du_dy = T(0.5) * (uyp_ - uym_);
    T du_dz = T(0.5) * (uzp_ - uzm_);
    T dv_dx = T(0.5) * (vxp - vxm);
    T dv_dy = T(0.5) * (vyp - vym);
    T dv_dz = T(0.5) * (vzp - vzm);
    T dw_dx = T(0.5) * (wxp - wxm);
    T dw_dy = T(0.5) * (wyp - wym);
    T dw_dz = T(0.5) * (wzp - wzm);
    T Sxx = du_dx, Syy = dv_dy, Szz = dw_dz;
    T Sxy = T(0.5) * (du_dy + dv_dx);
    T Sxz = T(0.5) * (du_dz + dw_dx);
    T Syz = T(0.5) * (dv_dz + dw_dy);
    T S = (Sxx * Sxx + Syy * Syy + Szz * Szz + T(2.0) * (Sxy * Sxy + Sxz * Sxz + Syz * Syz));
    T Smag = sqrt(T(2.0) * S + T(1e-6));
    T nu_t = c.cs2_les * Smag;
    if (nu_t > c.nu_t_max) nu_t = c.nu_t_max;
    if (near_wall) {
        T d = phi;
        T dmin = T(0.001) * c.dx;
        if (d < dmin) d = dmin;
        T lv = (T)det_log((double)(d / c.z0_phys));
        if (lv < T(2.0)) lv = T(2.0);
        T urx = u_here - uwall_x, ury = v_here - uwall_y;
        T sp = sqrt(urx * urx + ury * ury);
        if (sp > T(1e-6)) {
            T u_tau = sp * c.kappa / lv;
            T tau_w = u_tau * u_tau;
            T strain = sp / T(0.5);
            T nu_wm = tau_w / strain;
            T nu_tw = nu_wm - c.nu0;
            if (nu_tw < T(0)) nu_tw = T(0);
            if (nu_tw > nu_t) nu_t = nu_tw;
        }
    }
    T nu_eff = c.nu0 + nu_t;
    T tau_eff = T(0.5) + T(3.0) * nu_eff;
    if (tau_eff > T(2.0)) tau_eff = T(2.0);
    if (nu_t_out) *nu_t_out = nu_t;
    if (smag_out) *smag_out = Smag;
    return T(1.0) / tau_eff;
}

// second moments of f - feq in q order, zero terms skipped (exact: adding a
// product with a zero lattice component leaves the sum unchanged)
template <class T>
__device__ __forceinline__ void pi_neq(const T d[19], T& Pxx, T& Pyy, T& Pzz, T& Pxy, T& Pxz, T& Pyz) {
    Pxx = T(0);
    Pyy = T(0);
    Pzz = T(0);
    Pxy = T(0);
    Pxz = T(0);
    Pyz = T(0);
#pragma unroll
    for (int q = 0; q < 19; q++) {
        const int a = cxq(q), b = cyq(q), e = czq(q);
        if (a != 0) Pxx += d[q] * (T)a * (T)a;
        if (b != 0) Pyy += d[q] * (T)b * (T)b;
        if (e != 0) Pzz += d[q] * (T)e * (T)e;
        if (a * b != 0) Pxy += d[q] * (T)a * (T)b;
        if (a * e != 0) Pxz += d[q] * (T)a * (T)e;
        if (b * e != 0) Pyz += d[q] * (T)b * (T)e;
    }
}

// Guo forcing term (CODEX:3375-3383)
template <class T>
__device__ __forceinline__ T guo_force(int q, T om, T ux, T uy, T uz, T force_z) {
    const T cs2 = T(1.0) / T(3.0);
    T cxi = (T)cxq(q), cyi = (T)cyq(q), czi = (T)czq(q);
    T cdu = cxi * ux + cyi * uy + czi * uz;
    return wq<T>(q) * (T(1.0) - T(0.5) * om) * (((czi - uz) / cs2) + (cdu * czi / (cs2 * cs2))) * force_z;
}

// positivity limiter (CODEX:3391-3403)
template <class T>
__device__ __forceinline__ T limiter_alpha(const T fs[19], const T fe[19]) {
    T alpha = T(1.0);
    const T eps = T(1e-6);
#pragma unroll
    for (int q = 0; q < 19; q++)
        if (fs[q] < eps) {
            T den = fs[q] - fe[q];
            if (den < T(0)) {
                T cand = (eps - fe[q]) / den;
                if (cand < alpha) alpha = cand;
            }
        }
    if (alpha < T(0)) alpha = T(0);
    if (alpha > T(1)) alpha = T(1);
    return alpha;
}

// collide_regularized for one non-gas cell (CODEX:3353-3416); force_z = 0
// in air mode (the force term is then skipped exactly as the reference's
// ti.static branch does)
template <class T, bool FORCE>
__device__ __forceinline__ void collide_regularized(T f[19], T rho, T ux, T uy, T uz, T om, T force_z) {
    const T cs2 = T(1.0) / T(3.0);
    const T inv2 = T(1.0) / (T(2.0) * cs2 * cs2);
    T fe[19], d[19];
    T u2 = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        T cu = T(3.0) * ((T)cxq(q) * ux + (T)cyq(q) * uy + (T)czq(q) * uz);
        fe[q] = rho * wq<T>(q) * (T(1.0) + cu + T(0.5) * cu * cu - T(1.5) * u2);
        d[q] = f[q] - fe[q];
    }
    T Pxx, Pyy, Pzz, Pxy, Pxz, Pyz;
    pi_neq(d, Pxx, Pyy, Pzz, Pxy, Pxz, Pyz);
#pragma unroll
    for (int q = 0; q < 19; q++) {
        T cxi = (T)cxq(q), cyi = (T)cyq(q), czi = (T)czq(q);
        T Qxx = cxi * cxi - cs2, Qyy = cyi * cyi - cs2, Qzz = czi * czi - cs2;
        T Qxy = cxi * cyi, Qxz = cxi * czi, Qyz = cyi * czi;
        T freg = wq<T>(q) * inv2 * (Qxx * Pxx + Qyy * Pyy + Qzz * Pzz + T(2.0) * (Qxy * Pxy + Qxz * Pxz + Qyz * Pyz));
        T force_i = T(0);
        if (FORCE) force_i = guo_force<T>(q, om, ux, uy, uz, force_z);
        d[q] = fe[q] + (T(1.0) - om) * freg + force_i;  // f_star
    }
    T alpha = limiter_alpha(d, fe);
#pragma unroll
    for (int q = 0; q < 19; q++) f[q] = fe[q] + alpha * (d[q] - fe[q]);
}

// collide_KBC for one cell (CODEX:3172-3321); sponge handled by the caller
template <class T, bool FORCE>
__device__ __forceinline__ void collide_kbc(T f[19], T rho, T ux, T uy, T uz, T om, T force_z, bool backscatter, T c_bs, float xi_u) {
    const T cs2 = T(1.0) / T(3.0);
    const T eps = T(1e-6);
    T fe[19], fs[19];
    T u2 = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        T cu = T(3.0) * ((T)cxq(q) * ux + (T)cyq(q) * uy + (T)czq(q) * uz);
        fe[q] = rho * wq<T>(q) * (T(1.0) + cu + T(0.5) * cu * cu - T(1.5) * u2);
    }
#pragma unroll
    for (int q = 0; q < 19; q++) {
        T fv = f[q];
        T force_i = T(0);
        if (FORCE) force_i = guo_force<T>(q, om, ux, uy, uz, force_z);
        fs[q] = fv - om * (fv - fe[q]) + force_i;
    }
    T alpha = limiter_alpha(fs, fe);
#pragma unroll
    for (int q = 0; q < 19; q++) f[q] = fe[q] + alpha * (fs[q] - fe[q]);
    if (backscatter) {
        T d[19];
#pragma unroll
        for (int q = 0; q < 19; q++) d[q] = f[q] - fe[q];
        T Pxx, Pyy, Pzz, Pxy, Pxz, Pyz;
        pi_neq(d, Pxx, Pyy, Pzz, Pxy, Pxz, Pyz);
        T tr = (Pxx + Pyy + Pzz) / T(3.0);
        Pxx -= tr;
        Pyy -= tr;
        Pzz -= tr;
        T xi = T(2.0) * (T)xi_u - T(1.0);
        T scale = c_bs * xi;
        T inv2 = T(1.0) / (T(2.0) * cs2 * cs2);
        T gamma = T(1.0);
#pragma unroll
        for (int q = 0; q < 19; q++) {
            T cxi = (T)cxq(q), cyi = (T)cyq(q), czi = (T)czq(q);
            T Qxx = cxi * cxi - cs2, Qyy = cyi * cyi - cs2, Qzz = czi * czi - cs2;
            T Qxy = cxi * cyi, Qxz = cxi * czi, Qyz = cyi * czi;
            T contr = (Qxx * (scale * Pxx) + Qyy * (scale * Pyy) + Qzz * (scale * Pzz) +
                       T(2.0) * (Qxy * (scale * Pxy) + Qxz * (scale * Pxz) + Qyz * (scale * Pyz)));
            d[q] = wq<T>(q) * inv2 * contr;
            if (d[q] < T(0)) {
                T cand = (eps - f[q]) / d[q];
                if (cand < gamma) gamma = cand;
            }
        }
        if (gamma < T(0)) gamma = T(0);
        if (gamma > T(1)) gamma = T(1);
#pragma unroll
        for (int q = 0; q < 19; q++) f[q] = f[q] + gamma * d[q];
    }
}

// top sponge of collide_KBC (CODEX:3201-3208)
template <class T>
__device__ __forceinline__ void sponge_relax(T f[19], T rho, T ux, T uy, T uz) {
    const T om = T(1.0) / T(0.95);
    T u2 = ux * ux + uy * uy + uz * uz;
#pragma unroll
    for (int q = 0; q < 19; q++) {
        T cu = T(3.0) * ((T)cxq(q) * ux + (T)cyq(q) * uy + (T)czq(q) * uz);
        T fe = rho * wq<T>(q) * (T(1.0) + cu + T(0.5) * cu * cu - T(1.5) * u2);
        f[q] = f[q] - om * (f[q] - fe);
    }
}

}  // namespace lbm
