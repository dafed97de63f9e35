/*
 * lbm_oracle.c -- CPU restatement of the reference D3Q19 time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * library: it is built into oracle/liblbm_oracle.so and loaded only by the
 * tests, by __graft_entry__.smoke() and by bench.py's CPU-baseline leg.  The
 * product never links, loads or calls it.
 *
 * It restates, kernel by kernel and in the reference's order, the Taichi
 * kernels of LBM_3D_LES_air_over_waves_CODEX.py ("CODEX" below).  Fields are
 * kept whole (no fusion), exactly as the reference keeps them.  The file is
 * compiled twice: REAL=float (the reference's ti.f32 arithmetic, suffix _f32)
 * and REAL=double (the FP64 run, suffix _f64), always with
 * -ffp-contract=off so that every operation rounds once, in source order.
 *
 * Deviations from a literal reading, each forced by the reference:
 *  - sin/cos of the bed kinematics use det_sin/det_cos below (a fixed
 *    Cody-Waite + Taylor evaluation in double, rounded once); Taichi's
 *    ti.sin is backend-specific, so a fixed rule is needed for the bit-exact
 *    flag comparison with the device.
 *  - The racy VOF scatters (CODEX:1418-1575, 1660-1777, 1862-1994) read a
 *    snapshot of mass/fill/type taken when the kernel starts and add the
 *    scattered shares per receiving cell in a fixed order (own retained
 *    excess first, then q = 1..18 [, dist = 2..r]).  This is one of the
 *    schedules the reference admits; it conserves mass exactly as written.
 *  - The trapped-gas flood fill (CODEX:2021-2056) is iterated to its fixed
 *    point (exact 6-connected reachability) rather than stopped after
 *    Nx+Ny+Nz in-place sweeps; float atomic sums become FP64 sums.
 *  - The backscatter random number (CODEX:3280, ti.random) is replaced by a
 *    counter-based hash of (seed, step, cell).
 *  - New options: inflow/outflow x boundary, solid y walls, free-slip top,
 *    single-harmonic moving bed, settable C_s and tau (via the config).
 *  - FP16/BF16 population storage (new): the FP32 build applies the
 *    device's storage round trip wherever the device stores populations
 *    (store_round below); the FP64 build ignores the storage setting and is
 *    the full-precision run the reduced forms are measured against.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../include/lbm.h"
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef LBM_REAL
#define LBM_REAL float
#endif
typedef LBM_REAL REAL;
#define R(x) ((REAL)(x))
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#ifndef SFX
#define SFX _f32
#endif
#define API(name) CAT(name, SFX)

/* ---- D3Q19 tables (lbm_lattice.py:3-43) ------------------------------ */
static const int CX[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0};
static const int CY[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, -1, 1};
static const int CZ[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const int OPP[19] = {0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12, 17, 18, 15, 16};
static const double WD[19] = {1.0 / 3.0,  1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0, 1.0 / 18.0,
                              1.0 / 18.0, 1.0 / 18.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                              1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                              1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
/* z-mirror of each direction (free-slip top, new) */
static int MIRZ[19];

/* ---- deterministic sine / cosine ------------------------------------ */
/* Cody-Waite reduction by pi/2 (three-part constant) + Taylor series on
   |r| <= pi/4, evaluated in double without contraction. */
static double det_sincos_d(double x, int want_cos) {
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
    /* sin r = r - r^3/3! + ... + r^17/17! */
    double s = 1.0 / 355687428096000.0; /* 1/17! */
    s = -1.0 / 1307674368000.0 + r2 * s; /* 1/15! */
    s = 1.0 / 6227020800.0 + r2 * s;     /* 1/13! */
    s = -1.0 / 39916800.0 + r2 * s;      /* 1/11! */
    s = 1.0 / 362880.0 + r2 * s;         /* 1/9! */
    s = -1.0 / 5040.0 + r2 * s;          /* 1/7! */
    s = 1.0 / 120.0 + r2 * s;            /* 1/5! */
    s = -1.0 / 6.0 + r2 * s;             /* 1/3! */
    double sinr = r + (r * r2) * s;
    /* cos r = 1 - r^2/2! + ... + r^18/18! */
    double c = -1.0 / 6402373705728000.0; /* 1/18! */
    c = 1.0 / 20922789888000.0 + r2 * c; /* 1/16! */
    c = -1.0 / 87178291200.0 + r2 * c;     /* 1/14! */
    c = 1.0 / 479001600.0 + r2 * c;      /* 1/12! */
    c = -1.0 / 3628800.0 + r2 * c;         /* 1/10! */
    c = 1.0 / 40320.0 + r2 * c;          /* 1/8! */
    c = -1.0 / 720.0 + r2 * c;             /* 1/6! */
    c = 1.0 / 24.0 + r2 * c;             /* 1/4! */
    double cosr = (1.0 - 0.5 * r2) + (r2 * r2) * c;
    switch (quad) {
        case 0: return sinr;
        case 1: return cosr;
        case 2: return -sinr;
        default: return -cosr;
    }
}
/* natural logarithm: x = m 2^e with m in [sqrt(1/2), sqrt(2)), then
   log m = 2 atanh(s), s = (m-1)/(m+1), as an odd series to s^23, plus
   e ln 2 with ln 2 split in two parts; double, no contraction. */
/* exp(x) in double by Cody-Waite reduction by ln 2 (two-part constant) and a
   degree-14 Taylor polynomial on |r| <= ln2/2, scaled by 2^n; evaluated
   without contraction, so the device and the oracle agree bit for bit. */
static double det_exp_d(double x) {
    if (x != x) return x;
    if (x > 709.78) return HUGE_VAL;
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

static double det_log_d(double x) {
    if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
    if (isinf(x)) return x;
    int e;
    double m = frexp(x, &e); /* m in [0.5, 1) */
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
double API(orc_det_log)(double x) { return det_log_d(x); }
double API(orc_det_exp)(double x) { return det_exp_d(x); }
double API(orc_det_sin)(double x) { return det_sincos_d(x, 0); }
double API(orc_det_cos)(double x) { return det_sincos_d(x, 1); }
static REAL dsin(REAL x) { return (REAL)det_sincos_d((double)x, 0); }
static REAL dcos(REAL x) { return (REAL)det_sincos_d((double)x, 1); }

/* ---- state ------------------------------------------------------------ */
typedef struct {
    lbm_config_t c;
    int nx, ny, nz;
    size_t n, n2;
    int64_t step;
    REAL w[19];
    REAL *f, *fn;                  /* [19][n] */
    REAL *rho, *ux, *uy, *uz, *omega, *nu_t, *smag, *phi;
    int32_t *lo, *wt;              /* [19][n] */
    REAL *lf;                      /* [19][n] */
    int32_t *near, *obst, *obst_prev;
    REAL *water_h, *water_h_last, *u_wave, *v_wave, *w_wave; /* [ny][nx] */
    REAL *piecewise_bed_z;         /* [nx] */
    REAL *u_inlet_profile;         /* [nz] */
    REAL *inflow_ux;               /* [nz][ny] */
    /* water / VOF */
    REAL *fill, *mass, *mass_delta, *excess, *det_delta, *pref_head, *grav_frac;
    int32_t *type, *type_prev, *excess_dir, *ext_gas;
    REAL *fs_h, *fs_h_last;        /* [ny][nx] */
    REAL *fs_qx, *fs_qy;           /* [ny][nx] column volume fluxes (height tracking) */
    /* scratch for the deterministic VOF formulation */
    REAL *s_m, *s_a, *s_b, *s_c, *s_d, *s_e, *s_mv;
    int32_t *s_i;
    lbm_counters_t cnt;
} O;

#define IDX(k, j, i) (((size_t)(k) * (size_t)o->ny + (size_t)(j)) * (size_t)o->nx + (size_t)(i))
#define ID2(j, i) ((size_t)(j) * (size_t)o->nx + (size_t)(i))
#define F(q, id) (o->f[(size_t)(q) * o->n + (id)])
#define FN(q, id) (o->fn[(size_t)(q) * o->n + (id)])
#define LO(q, id) (o->lo[(size_t)(q) * o->n + (id)])
#define LF(q, id) (o->lf[(size_t)(q) * o->n + (id)])
#define WT(q, id) (o->wt[(size_t)(q) * o->n + (id)])

static int pmod(int a, int n) {
    int r = a % n;
    return r < 0 ? r + n : r;
}
static REAL smoothstep01(REAL x) { return x * x * (R(3.0) - R(2.0) * x); } /* CODEX:654-657 */

static int bed_is_moving(const O *o) {
    return o->c.bed_profile == LBM_BED_MOVING_SINE || o->c.bed_profile == LBM_BED_MOVING_SINE1;
}
static int is_water(const O *o) { return o->c.fs_enabled == 1; }
static int is_vof(const O *o) { return o->c.fs_enabled == 1 && o->c.fs_tracking == LBM_FST_VOF; }
static int use_hbp(const O *o) {
    return o->c.fs_enabled == 1 && (o->c.pressure_formulation == LBM_PF_HYDROSTATIC_BALANCED ||
                                    o->c.pressure_formulation == LBM_PF_VOF_COMPONENT_BALANCED);
}
static int use_cbp(const O *o) {
    return o->c.fs_enabled == 1 && o->c.pressure_formulation == LBM_PF_VOF_COMPONENT_BALANCED;
}

/* calculate_feq (CODEX:623-643) */
static REAL feq(const O *o, REAL rho, REAL ux, REAL uy, REAL uz, int q) {
    REAL u2 = ux * ux + uy * uy + uz * uz;
    REAL cu = R(3.0) * ((REAL)CX[q] * ux + (REAL)CY[q] * uy + (REAL)CZ[q] * uz);
    return rho * o->w[q] * (R(1.0) + cu + R(0.5) * cu * cu - R(1.5) * u2);
}

/* neighbour lookup used by every VOF loop: x periodic or bounded, y periodic
   (or bounded with solid y walls), z bounded.  CODEX:1322-1333 pattern. */
static int nbr(const O *o, int i, int j, int k, int di, int dj, int dk, int *ni, int *nj, int *nk) {
    int a = i + di, b = j + dj, c = k + dk;
    int valid = 1;
    if (o->c.x_boundary == LBM_XB_PERIODIC)
        a = pmod(a, o->nx);
    else if (a < 0 || a >= o->nx)
        valid = 0;
    if (o->c.y_boundary == LBM_YB_PERIODIC)
        b = pmod(b, o->ny);
    else if (b < 0 || b >= o->ny)
        valid = 0;
    if (c < 0 || c >= o->nz) valid = 0;
    *ni = a;
    *nj = b;
    *nk = c;
    return valid;
}

/* ---- geometry --------------------------------------------------------- */

/* update_wave_bed_and_velocities (CODEX:1082-1162); moving_sine1 is the new
   single-harmonic bed eta = A sin(k (x - c t)). */
static void bed_update(O *o, REAL t) {
    const lbm_config_t *c = &o->c;
    REAL dx = R(c->dx);
    REAL kw = R(2.0) * R(3.141592653589793) / R(c->bed_wavelength_m);
    REAL cs = R(c->wave_speed);
    REAL ramp = R(2.0) * R(c->wave_period);
    REAL amp_now = R(0.0);
    if (c->bed_profile == LBM_BED_MOVING_SINE)
        amp_now = t < ramp ? R(c->bed_amp_m) * smoothstep01(t / ramp) : R(c->bed_amp_m);
    REAL c1 = R(c->bed_phase_speed_mps);
    REAL dtp = R(c->dt_phys);
    for (int j = 0; j < o->ny; j++)
        for (int i = 0; i < o->nx; i++) {
            size_t id = ID2(j, i);
            REAL x0 = (REAL)i * dx;
            o->water_h_last[id] = o->water_h[id];
            REAL eta1 = R(0), eta2 = R(0), eta = R(0), phase = R(0), phase2 = R(0);
            if (c->bed_profile == LBM_BED_MOVING_SINE) {
                REAL x_phase = x0 - cs * t;
                REAL x_phase2 = x0 - R(0.5) * cs * t;
                phase = R(0.5) * kw * x_phase;
                phase2 = kw * x_phase2;
                eta1 = amp_now * dsin(phase);
                eta2 = amp_now * dsin(phase2);
                eta = eta1 + eta2;
            } else if (c->bed_profile == LBM_BED_MOVING_SINE1) {
                phase = kw * (x0 - c1 * t);
                eta1 = R(c->bed_amp_m) * dsin(phase);
                eta = eta1;
            }
            REAL bed_h = R(c->offset_m) + eta;
            if (c->bed_profile == LBM_BED_PIECEWISE) bed_h = o->piecewise_bed_z[i];
            o->water_h[id] = bed_h;
            REAL u_s = R(0), v_s = R(0), w_s = R(0);
            REAL h0 = R(c->depth_swl_m);
            if (c->bed_profile == LBM_BED_MOVING_SINE) {
                REAL eta_x1 = amp_now * R(0.5) * kw * dcos(phase);
                REAL eta_x2 = amp_now * kw * dcos(phase2);
                REAL eta_x = eta_x1 + eta_x2;
                REAL eta_y = R(0);
                REAL eta_t = (o->water_h[id] - o->water_h_last[id]) / dtp;
                u_s = cs * eta1 / (h0 + eta1) + R(0.5) * cs * eta2 / (h0 + eta2);
                w_s = eta_t + u_s * eta_x + v_s * eta_y;
            } else if (c->bed_profile == LBM_BED_MOVING_SINE1) {
                REAL eta_x = R(c->bed_amp_m) * kw * dcos(phase);
                REAL eta_t = (o->water_h[id] - o->water_h_last[id]) / dtp;
                u_s = c1 * eta1 / (h0 + eta1);
                w_s = eta_t + u_s * eta_x;
            }
            REAL vs = R(c->vel_scale);
            o->u_wave[id] = u_s / vs;
            o->v_wave[id] = v_s / vs;
            o->w_wave[id] = w_s / vs;
        }
}

/* periodic bilinear sample of a cell-centred 2-D field (CODEX:2536-2569,
   937-988).  The wrapped index is folded back into range when rounding puts
   it exactly on the upper bound. */
static REAL sample2d(const O *o, const REAL *H, REAL xq, REAL yq) {
    REAL dx = R(o->c.dx);
    REAL Lx = (REAL)o->nx * dx, Ly = (REAL)o->ny * dx;
    REAL Nxf = (REAL)o->nx, Nyf = (REAL)o->ny;
    REAL xw = xq - floor(xq / Lx) * Lx;
    REAL yw = yq - floor(yq / Ly) * Ly;
    REAL u = xw / dx - R(0.5);
    REAL v = yw / dx - R(0.5);
    u = u - floor(u / Nxf) * Nxf;
    v = v - floor(v / Nyf) * Nyf;
    int i0 = (int)floor(u), j0 = (int)floor(v);
    REAL fx = u - (REAL)i0, fy = v - (REAL)j0;
    if (i0 >= o->nx) i0 -= o->nx;
    if (j0 >= o->ny) j0 -= o->ny;
    int i1 = (i0 + 1) % o->nx, j1 = (j0 + 1) % o->ny;
    REAL h00 = H[ID2(j0, i0)], h10 = H[ID2(j0, i1)], h01 = H[ID2(j1, i0)], h11 = H[ID2(j1, i1)];
    REAL hx0 = h00 * (R(1.0) - fx) + h10 * fx;
    REAL hx1 = h01 * (R(1.0) - fx) + h11 * fx;
    return hx0 * (R(1.0) - fy) + hx1 * fy;
}

/* cube_sdf_at_point (CODEX:897-934): signed distance to the box obstacle,
   offsets folded into the periodic domain (x only when x is periodic, y
   always), rotated by yaw about z then pitch about y. */
static REAL cube_sdf(const O *o, REAL xp, REAL yp, REAL zp) {
    const lbm_config_t *c = &o->c;
    const REAL lx = (REAL)o->nx * R(c->dx), ly = (REAL)o->ny * R(c->dx);
    REAL dxp = xp - R(c->cube_center[0]);
    REAL dyp = yp - R(c->cube_center[1]);
    if (c->x_boundary == LBM_XB_PERIODIC) {
        if (dxp > R(0.5) * lx) dxp -= lx;
        if (dxp < R(-0.5) * lx) dxp += lx;
    }
    if (dyp > R(0.5) * ly) dyp -= ly;
    if (dyp < R(-0.5) * ly) dyp += ly;
    const REAL ycos = R(c->cube_rot[0]), ysin = R(c->cube_rot[1]), pcos = R(c->cube_rot[2]), psin = R(c->cube_rot[3]);
    REAL x_local = ycos * dxp + ysin * dyp;
    REAL y_local = -ysin * dxp + ycos * dyp;
    REAL z1 = zp - R(c->cube_center[2]);
    REAL x_pitch = pcos * x_local - psin * z1;
    REAL z_local = psin * x_local + pcos * z1;
    REAL qx = (REAL)fabs((double)x_pitch) - R(c->cube_half[0]);
    REAL qy = (REAL)fabs((double)y_local) - R(c->cube_half[1]);
    REAL qz = (REAL)fabs((double)z_local) - R(c->cube_half[2]);
    REAL ox = qx > R(0) ? qx : R(0), oy = qy > R(0) ? qy : R(0), oz = qz > R(0) ? qz : R(0);
    REAL outside = (REAL)sqrt((double)(ox * ox + oy * oy + oz * oz));
    REAL m = qy > qz ? qy : qz;
    m = qx > m ? qx : m;
    REAL inside = m < R(0) ? m : R(0);
    return outside + inside;
}

/* compute_phi_slope_corrected (CODEX:2432-2490), bed, box and x-wall terms. */
static void compute_phi(O *o) {
    const lbm_config_t *c = &o->c;
    REAL dx = R(c->dx);
    for (int j = 0; j < o->ny; j++)
        for (int i = 0; i < o->nx; i++) {
            int ip = (i + 1) % o->nx, im = (i - 1 + o->nx) % o->nx;
            int jp = (j + 1) % o->ny, jm = (j - 1 + o->ny) % o->ny;
            REAL dh_dx = (o->water_h[ID2(j, ip)] - o->water_h[ID2(j, im)]) * (R(0.5) / dx);
            if (c->x_boundary != LBM_XB_PERIODIC) {
                if (i == 0)
                    dh_dx = (o->water_h[ID2(j, i + 1)] - o->water_h[ID2(j, i)]) / dx;
                else if (i == o->nx - 1)
                    dh_dx = (o->water_h[ID2(j, i)] - o->water_h[ID2(j, i - 1)]) / dx;
            }
            REAL dh_dy = (o->water_h[ID2(jp, i)] - o->water_h[ID2(jm, i)]) * (R(0.5) / dx);
            REAL denom = (REAL)sqrt(R(1.0) + dh_dx * dh_dx + dh_dy * dh_dy);
            if (denom < R(1e-6)) denom = R(1.0);
            REAL h_loc = o->water_h[ID2(j, i)];
            for (int k = 0; k < o->nz; k++) {
                REAL zc = ((REAL)k + R(0.5)) * dx;
                REAL xc = ((REAL)i + R(0.5)) * dx;
                REAL phi_val = R(1.0e6);
                if (c->include_bed) phi_val = (zc - h_loc) / denom;
                if (c->include_cube) {
                    REAL cp = cube_sdf(o, xc, ((REAL)j + R(0.5)) * dx, zc);
                    if (cp < phi_val) phi_val = cp;
                }
                if (c->x_boundary == LBM_XB_SOLID) {
                    REAL a = xc, b = (REAL)o->nx * dx - xc;
                    REAL xw = a < b ? a : b;
                    if (xw < phi_val) phi_val = xw;
                }
                o->phi[IDX(k, j, i)] = phi_val;
            }
        }
}

/* sample_phi_periodic (CODEX:661-707): trilinear sample of phi, periodic
   bilinear in x / y (the bed sampling's indices), clamped in z */
static REAL sample_phi(const O *o, REAL xp, REAL yp, REAL zp) {
    const REAL dx = R(o->c.dx);
    const REAL Lx = (REAL)o->nx * dx, Ly = (REAL)o->ny * dx;
    const REAL Nxf = (REAL)o->nx, Nyf = (REAL)o->ny;
    REAL xw = xp - (REAL)floor((double)(xp / Lx)) * Lx;
    REAL yw = yp - (REAL)floor((double)(yp / Ly)) * Ly;
    REAL u = xw / dx - R(0.5);
    REAL v = yw / dx - R(0.5);
    u = u - (REAL)floor((double)(u / Nxf)) * Nxf;
    v = v - (REAL)floor((double)(v / Nyf)) * Nyf;
    int i0 = (int)floor((double)u), j0 = (int)floor((double)v);
    REAL fx = u - (REAL)i0, fy = v - (REAL)j0;
    if (i0 >= o->nx) i0 -= o->nx;
    if (j0 >= o->ny) j0 -= o->ny;
    int i1 = (i0 + 1) % o->nx, j1 = (j0 + 1) % o->ny;
    REAL r = zp / dx - R(0.5);
    int k0 = (int)floor((double)r);
    REAL fz = r - (REAL)k0;
    int k1 = k0 + 1;
    if (k0 < 0) {
        k0 = 0;
        k1 = 0;
        fz = R(0);
    }
    if (k1 >= o->nz) {
        k0 = o->nz - 1;
        k1 = o->nz - 1;
        fz = R(0);
    }
    const REAL *ph = o->phi;
    REAL p00_0 = ph[IDX(k0, j0, i0)] * (R(1.0) - fx) + ph[IDX(k0, j0, i1)] * fx;
    REAL p01_0 = ph[IDX(k0, j1, i0)] * (R(1.0) - fx) + ph[IDX(k0, j1, i1)] * fx;
    REAL pxy_0 = p00_0 * (R(1.0) - fy) + p01_0 * fy;
    REAL p00_1 = ph[IDX(k1, j0, i0)] * (R(1.0) - fx) + ph[IDX(k1, j0, i1)] * fx;
    REAL p01_1 = ph[IDX(k1, j1, i0)] * (R(1.0) - fx) + ph[IDX(k1, j1, i1)] * fx;
    REAL pxy_1 = p00_1 * (R(1.0) - fy) + p01_1 * fy;
    return pxy_0 * (R(1.0) - fz) + pxy_1 * fz;
}

/* build_lattice_open_from_free_surface_periodic (CODEX:2492-2686): the
   open fraction from the obstacle gap ("vertical") or from phi sampled at
   the link endpoint ("phi_fraction", "phi"), the open flag from the gap or
   (mode "phi") from the same phi sample; bed and box obstacles (the nearer one sets the gap and the
   wall type); solid y walls mirror the x-wall override. */
static void build_lattice_open(O *o) {
    const lbm_config_t *c = &o->c;
    REAL dx = R(c->dx);
    int solid_x = c->x_boundary == LBM_XB_SOLID;
    int solid_y = c->y_boundary == LBM_YB_SOLID;
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                REAL xc = ((REAL)i + R(0.5)) * dx, yc = ((REAL)j + R(0.5)) * dx,
                     zc = ((REAL)k + R(0.5)) * dx;
                int lo[19];
                REAL lf[19];
                int wt[19];
                REAL eta_c = sample2d(o, o->water_h, xc, yc);
                REAL gap_c = zc - eta_c;
                REAL ggap_c = R(1.0e6);
                int wtc = LBM_WALL_NONE;
                if (c->include_bed) {
                    ggap_c = gap_c;
                    wtc = LBM_WALL_BED;
                }
                if (c->include_cube) {
                    REAL cg = cube_sdf(o, xc, yc, zc);
                    if (cg < ggap_c) {
                        ggap_c = cg;
                        wtc = LBM_WALL_CUBE;
                    }
                }
                REAL sgap_c = c->phi_fraction ? o->phi[id] : ggap_c;
                REAL frac_c = R(0.5) + sgap_c / dx;
                if (frac_c < R(0)) frac_c = R(0);
                if (frac_c > R(1)) frac_c = R(1);
                lf[0] = frac_c;
                wt[0] = wtc;
                for (int q = 1; q < 19; q++) {
                    REAL xq = xc + R(0.5) * dx * (REAL)CX[q];
                    REAL yq = yc + R(0.5) * dx * (REAL)CY[q];
                    REAL zq = zc + R(0.5) * dx * (REAL)CZ[q];
                    REAL eta = sample2d(o, o->water_h, xq, yq);
                    REAL gap = zq - eta;
                    REAL ggap = R(1.0e6);
                    int wq = LBM_WALL_NONE;
                    if (c->include_bed) {
                        ggap = gap;
                        wq = LBM_WALL_BED;
                    }
                    if (c->include_cube) {
                        REAL cg = cube_sdf(o, xq, yq, zq);
                        if (cg < ggap) {
                            ggap = cg;
                            wq = LBM_WALL_CUBE;
                        }
                    }
                    REAL sgap = c->phi_fraction ? sample_phi(o, xq, yq, zq) : ggap;
                    REAL frac = R(0.5) + sgap / dx;
                    if (frac < R(0)) frac = R(0);
                    if (frac > R(1)) frac = R(1);
                    lf[q] = frac;
                    wt[q] = wq;
                    REAL flag_gap = c->phi_flags ? sgap : ggap;
                    lo[q] = (flag_gap < R(0)) ? 0 : 1;
                    if (solid_x) {
                        if (i == 0 && CX[q] > 0) {
                            lo[q] = 0;
                            lf[q] = R(0);
                            wt[q] = LBM_WALL_XMIN;
                        } else if (i == o->nx - 1 && CX[q] < 0) {
                            lo[q] = 0;
                            lf[q] = R(0);
                            wt[q] = LBM_WALL_XMAX;
                        }
                    }
                    if (solid_y) {
                        if (j == 0 && CY[q] > 0) {
                            lo[q] = 0;
                            lf[q] = R(0);
                            wt[q] = LBM_WALL_YMIN;
                        } else if (j == o->ny - 1 && CY[q] < 0) {
                            lo[q] = 0;
                            lf[q] = R(0);
                            wt[q] = LBM_WALL_YMAX;
                        }
                    }
                }
                lo[0] = 0;
                for (int q = 1; q < 19; q++)
                    if (lo[q] == 1) lo[0] = 1;
                int near = 0;
                for (int q = 1; q < 19; q++)
                    if (lo[q] == 0) near = 1;
                if (lo[0] == 0)
                    for (int q = 1; q < 19; q++) lf[q] = R(0);
                o->near[id] = near;
                for (int q = 0; q < 19; q++) {
                    LO(q, id) = lo[q];
                    LF(q, id) = lf[q];
                    WT(q, id) = wt[q];
                }
            }
}

/* wall velocity of link q at cell (k,j,i): link sampling (CODEX:937-988) or
   the cell value. */
static void wall_velocity(const O *o, int j, int i, int q, REAL *uw) {
    REAL dx = R(o->c.dx);
    if (o->c.wall_velocity_link) {
        REAL xc = ((REAL)i + R(0.5)) * dx, yc = ((REAL)j + R(0.5)) * dx;
        REAL xq = xc + R(0.5) * dx * (REAL)CX[q];
        REAL yq = yc + R(0.5) * dx * (REAL)CY[q];
        uw[0] = sample2d(o, o->u_wave, xq, yq);
        uw[1] = sample2d(o, o->v_wave, xq, yq);
        uw[2] = sample2d(o, o->w_wave, xq, yq);
    } else {
        uw[0] = o->u_wave[ID2(j, i)];
        uw[1] = o->v_wave[ID2(j, i)];
        uw[2] = o->w_wave[ID2(j, i)];
    }
}

/* ---- pressure reference (CODEX:710-740, 646-651, 821-867) ----------- */
static void update_pressure_reference(O *o) {
    const lbm_config_t *c = &o->c;
    REAL dx = R(c->dx);
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                o->pref_head[id] = R(0);
                o->grav_frac[id] = R(0);
                if (is_water(o) && LO(0, id) == 1 && o->type[id] != LBM_FS_GAS) {
                    REAL zc = ((REAL)k + R(0.5)) * dx;
                    if (c->pressure_formulation == LBM_PF_HYDROSTATIC_BALANCED) {
                        o->pref_head[id] = R(c->free_surface_base_m) - zc;
                    } else if (c->pressure_formulation == LBM_PF_VOF_COMPONENT_BALANCED) {
                        if (zc < R(c->free_surface_base_m)) {
                            o->pref_head[id] = R(c->free_surface_base_m) - zc;
                        } else {
                            o->grav_frac[id] = R(1);
                        }
                    } else {
                        o->grav_frac[id] = R(1);
                    }
                }
            }
}

static REAL local_g(const O *o, size_t id) {
    if (use_cbp(o)) return R(o->c.g_lb) * o->grav_frac[id];
    return R(o->c.g_body_lb);
}

static REAL fs_density_cell(const O *o, int k, int j, int i) {
    const lbm_config_t *c = &o->c;
    REAL dx = R(c->dx);
    REAL head = R(0);
    size_t id = IDX(k, j, i);
    if (use_cbp(o)) {
        REAL zc = ((REAL)k + R(0.5)) * dx;
        head = o->grav_frac[id] < R(0.5) ? zc - R(c->free_surface_base_m) : R(0);
    } else if (use_hbp(o)) {
        head = o->fs_h[ID2(j, i)] - R(c->free_surface_base_m);
    } else {
        REAL zc = ((REAL)k + R(0.5)) * dx;
        head = o->fs_h[ID2(j, i)] - zc;
    }
    if (head < R(0) && !use_hbp(o)) head = R(0);
    return R(c->rho0) + R(c->g_lb) * (head / dx) / (R(1.0) / R(3.0));
}

static REAL fs_density_link(const O *o, int k, int j, int i, int q) {
    const lbm_config_t *c = &o->c;
    REAL dx = R(c->dx);
    size_t id = IDX(k, j, i);
    if (use_cbp(o)) {
        REAL z_if = ((REAL)k + R(0.5)) * dx - R(0.5) * (REAL)CZ[q] * dx;
        REAL head = R(0);
        if (o->grav_frac[id] < R(0.5)) head += z_if - R(c->free_surface_base_m);
        return R(c->rho0) + R(c->g_lb) * (head / dx) / (R(1.0) / R(3.0));
    } else if (use_hbp(o) && c->fs_tracking == LBM_FST_VOF) {
        REAL z_if = ((REAL)k + R(0.5)) * dx - R(0.5) * (REAL)CZ[q] * dx;
        REAL head = z_if - R(c->free_surface_base_m);
        return R(c->rho0) + R(c->g_lb) * (head / dx) / (R(1.0) / R(3.0));
    }
    return fs_density_cell(o, k, j, i);
}

/* vof_cell_liquid_fraction (CODEX:870-884) */
static REAL liquid_fraction(const O *o, size_t id) {
    REAL a = R(1);
    if (is_vof(o) && LO(0, id) == 1) {
        if (o->type[id] == LBM_FS_GAS)
            a = R(0);
        else if (o->type[id] == LBM_FS_INTERFACE) {
            a = o->fill[id];
            if (a < R(0)) a = R(0);
            if (a > R(1)) a = R(1);
        }
    }
    return a;
}

/* ---- step kernels ------------------------------------------------------ */

/* the 16-bit codec of the run: shift s_q, scale 2^k and rounding (on = 0
   for the FP64 build and for 32/64-bit storage, which round nothing) */
typedef struct {
    int on, sr, half;
    float sh[19], scale, inv;
} Codec16;
static void codec16(const O *o, Codec16 *k);
static REAL round16(const Codec16 *k, REAL v, int q, uint32_t base);
static int staged_rounds(const O *o);
static uint32_t bits_of(REAL v);
static uint32_t dither_base(uint32_t step, uint32_t cell);
enum { SALT_STREAM = 0x40000000u, SALT_REFILL = 0x20000000u, SALT_COND = 0x10000000u, SALT_INIT = 0x08000000u };

/* update_bed_populations_and_reinit (CODEX:2878-2940) */
static void bed_populations(O *o) {
    const lbm_config_t *c = &o->c;
    int air = c->physics_mode == LBM_PHYS_AIR;
    int moving = bed_is_moving(o);
    Codec16 cd;
    codec16(o, &cd);
    const int rnd = staged_rounds(o);
    const uint32_t salt_step = (uint32_t)o->step ^ SALT_COND;
    for (size_t id = 0; id < o->n; id++) {
        o->obst_prev[id] = o->obst[id];
        o->obst[id] = LO(0, id) == 0 ? 1 : 0;
    }
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                for (int q = 0; q < 19; q++) {
                    if (LO(q, id) != 0) continue;
                    REAL frac = LF(q, id);
                    int wtq = WT(q, id);
                    REAL uw[3] = {R(0), R(0), R(0)};
                    if (wtq == LBM_WALL_BED && moving) {
                        wall_velocity(o, j, i, q, uw);
                    } else if (air) {
                        if (wtq == LBM_WALL_BED) wall_velocity(o, j, i, q, uw);
                    } else {
                        continue;
                    }
                    REAL fe = feq(o, R(c->rho0), uw[0], uw[1], uw[2], q);
                    REAL nv = frac * F(q, id) + (R(1.0) - frac) * fe;
                    if (rnd && bits_of(nv) != bits_of(F(q, id))) nv = round16(&cd, nv, q, dither_base(salt_step, (uint32_t)id));
                    F(q, id) = nv;
                }
            }
}

/* macro_step (CODEX:2943-3008) */
static void macro(O *o) {
    const lbm_config_t *c = &o->c;
#pragma omp parallel for schedule(static)
    for (size_t id = 0; id < o->n; id++) {
        REAL r = R(0);
        for (int q = 0; q < 19; q++) r += F(q, id);
        o->rho[id] = r;
    }
#pragma omp parallel for schedule(static)
    for (size_t id = 0; id < o->n; id++) {
        REAL jx = R(0), jy = R(0), jz = R(0);
        for (int q = 0; q < 19; q++) {
            REAL fv = F(q, id);
            jx += fv * (REAL)CX[q];
            jy += fv * (REAL)CY[q];
            jz += fv * (REAL)CZ[q];
        }
        if (is_water(o)) {
            if (o->type[id] == LBM_FS_GAS) {
                o->rho[id] = R(c->rho0);
                o->ux[id] = o->uy[id] = o->uz[id] = R(0);
            } else {
                REAL force_z = -o->rho[id] * local_g(o, id);
                REAL inv = R(1.0) / o->rho[id];
                o->ux[id] = jx * inv;
                o->uy[id] = jy * inv;
                o->uz[id] = (jz + R(0.5) * force_z) * inv;
            }
        } else {
            REAL inv = R(1.0) / o->rho[id];
            o->ux[id] = jx * inv;
            o->uy[id] = jy * inv;
            o->uz[id] = jz * inv;
        }
    }
    if (c->enable_top_drive) {
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(o->nz - 1, j, i);
                o->ux[id] = R(c->u_top_lb);
                o->uy[id] = R(0);
                o->uz[id] = R(0);
            }
    }
}

/* compute_LES (CODEX:3011-3141) */
static void les(O *o) {
    const lbm_config_t *c = &o->c;
    REAL dx = R(c->dx);
    REAL csd = R(c->cs_les) * R(c->delta_les);
    REAL cs2 = csd * csd;
    REAL nu0 = R(c->nu0);
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                int active = LO(0, id) == 1;
                if (is_water(o) && o->type[id] == LBM_FS_GAS) active = 0;
                if (!active) continue;
                int ip = (i + 1) % o->nx, im = (i + o->nx - 1) % o->nx;
                int jp = (j + 1) % o->ny, jm = (j + o->ny - 1) % o->ny;
                int kp = k + 1 >= o->nz ? o->nz - 1 : k + 1;
                int km = k - 1 < 0 ? 0 : k - 1;
                size_t xp = IDX(k, j, ip), xm = IDX(k, j, im), yp = IDX(k, jp, i), ym = IDX(k, jm, i),
                       zp = IDX(kp, j, i), zm = IDX(km, j, i);
                REAL du_dx = R(0.5) * (o->ux[xp] - o->ux[xm]);
                REAL du_dy = R(0.5) * (o->ux[yp] - o->ux[ym]);
                REAL du_dz = R(0.5) * (o->ux[zp] - o->ux[zm]);
                REAL dv_dx = R(0.5) * (o->uy[xp] - o->uy[xm]);
                REAL dv_dy = R(0.5) * (o->uy[yp] - o->uy[ym]);
                REAL dv_dz = R(0.5) * (o->uy[zp] - o->uy[zm]);
                REAL dw_dx = R(0.5) * (o->uz[xp] - o->uz[xm]);
                REAL dw_dy = R(0.5) * (o->uz[yp] - o->uz[ym]);
                REAL dw_dz = R(0.5) * (o->uz[zp] - o->uz[zm]);
                REAL Sxx = du_dx, Syy = dv_dy, Szz = dw_dz;
                REAL Sxy = R(0.5) * (du_dy + dv_dx);
                REAL Sxz = R(0.5) * (du_dz + dw_dx);
                REAL Syz = R(0.5) * (dv_dz + dw_dy);
                REAL S = (Sxx * Sxx + Syy * Syy + Szz * Szz + R(2.0) * (Sxy * Sxy + Sxz * Sxz + Syz * Syz));
                REAL Smag = (REAL)sqrt(R(2.0) * S + R(1e-6));
                REAL nu_t = cs2 * Smag;
                if (nu_t > R(c->nu_t_max)) nu_t = R(c->nu_t_max);
                if (k > 0 && LO(0, IDX(k - 1, j, i)) == 0) {
                    REAL d = o->phi[id];
                    REAL dmin = R(0.001) * dx;
                    if (d < dmin) d = dmin;
                    REAL lv = (REAL)det_log_d((double)(d / R(c->z0_phys)));
                    if (lv < R(2.0)) lv = R(2.0);
                    int wb = WT(0, IDX(k - 1, j, i));
                    REAL uwx = R(0), uwy = R(0);
                    if (wb == LBM_WALL_BED) {
                        uwx = o->u_wave[ID2(j, i)];
                        uwy = o->v_wave[ID2(j, i)];
                    }
                    REAL urx = o->ux[id] - uwx, ury = o->uy[id] - uwy;
                    REAL sp = (REAL)sqrt(urx * urx + ury * ury);
                    if (sp > R(1e-6)) {
                        REAL u_tau = sp * R(c->kappa) / lv;
                        REAL tau_w = u_tau * u_tau;
                        REAL strain = sp / R(0.5);
                        REAL nu_wm = tau_w / strain;
                        REAL nu_tw = nu_wm - nu0;
                        if (nu_tw < R(0)) nu_tw = R(0);
                        if (nu_tw > nu_t) nu_t = nu_tw;
                    }
                }
                REAL nu_eff = nu0 + nu_t;
                REAL tau_eff = R(0.5) + R(3.0) * nu_eff;
                if (tau_eff > R(2.0)) tau_eff = R(2.0);
                o->nu_t[id] = nu_t;
                o->smag[id] = Smag;
                o->omega[id] = R(1.0) / tau_eff;
            }
}

/* counter-based uniform [0,1) for backscatter (new; replaces ti.random) */
static uint32_t fmix32(uint32_t h) {
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}
static float bs_uniform(uint32_t seed, uint32_t step, uint32_t cell) {
    uint32_t h = fmix32(seed ^ fmix32(cell + fmix32(step * 0x9E3779B9u)));
    return (float)(h >> 8) * (1.0f / 16777216.0f);
}

/* ---- reduced-precision population storage (new) ----------------------
 * The device stores h_q = round((f_q - s_q) 2^k) as binary16 or bfloat16
 * and reads f_q = s_q + h_q 2^-k, s_q = feq_q(1, u_ref).  The FP32 oracle
 * applies the same round trip wherever the device stores populations
 * (initialisation, upload of F, after every collision; on the staged path
 * also after the stream and where the conditioning, the refill and the
 * initial free-surface projection write), so the FP16/BF16 device run is
 * checked against it like the FP32 one.  The roundings are
 * restated from the IEEE 754 binary16 and bfloat16 formats:
 *  - nearest: round to nearest, ties to even (half: 10 fraction bits above
 *    2^-14, quantum 2^-24 below; overflow to infinity);
 *  - stochastic: a 32-bit dither r of (step, cell, q) is added to the
 *    magnitude's bits below the kept fraction bits, which are then
 *    truncated (half: r >> 19 onto 13 dropped bits, then the nearest-even
 *    conversion, exact except in half's subnormal range; bfloat16: r >> 16
 *    onto 16 dropped bits). */
static float half_rne(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    uint32_t a = v.u & 0x7fffffffu, sg = v.u & 0x80000000u;
    if (a >= 0x7f800000u) return x;
    if (a < 0x38800000u) return rintf(x * 16777216.0f) * (1.0f / 16777216.0f);
    a += 0xfffu + ((a >> 13) & 1u);
    a &= ~0x1fffu;
    if (a >= 0x47800000u) a = 0x7f800000u;
    v.u = a | sg;
    return v.f;
}
/* Stochastic rounding to binary16 with the 32-bit random word r: the value
   is rounded up to the next binary16 number iff the 13 bits of r above bit
   19 plus the first 13 bits of the value below its binary16 ulp carry into
   the ulp, i.e. with probability (distance to the lower neighbour) / ulp in
   steps of 2^-13.  In the normal range that is "add r >> 19 to the low 13
   fraction bits of the float and truncate" (overflow to infinity at 2^16);
   in the subnormal range the ulp is 2^-24 and the value's bits below it are
   counted in units of 2^-37.  This is the rounding of the gfx950
   instruction v_cvt_sr_f16_f32 (pinned by tests/golden/sr_hw_vectors.npz). */
static float half_sr(float x, uint32_t r) {
    union { float f; uint32_t u; } v;
    v.f = x;
    uint32_t a = v.u & 0x7fffffffu, sg = v.u & 0x80000000u;
    if (a >= 0x7f800000u) return x;
    if (a < 0x38800000u) {
        uint32_t t = (uint32_t)floor((double)fabsf(x) * 137438953472.0); /* |x| / 2^-37, exact, < 2^23 */
        uint32_t qv = (t + (r >> 19)) >> 13;                               /* units of 2^-24, <= 1024 */
        v.f = (float)qv * (1.0f / 16777216.0f);
        v.u |= sg;
        return v.f;
    }
    a = (a + (r >> 19)) & ~0x1fffu;
    if (a >= 0x47800000u) a = 0x7f800000u;
    v.u = a | sg;
    return v.f;
}
static float bf16_rne(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    if ((v.u & 0x7fffffffu) > 0x7f800000u) {
        v.u = ((v.u >> 16) | 0x40u) << 16;
        return v.f;
    }
    v.u += 0x7fffu + ((v.u >> 16) & 1u);
    v.u &= 0xffff0000u;
    return v.f;
}
static float bf16_sr(float x, uint32_t r) {
    union { float f; uint32_t u; } v;
    v.f = x;
    if ((v.u & 0x7fffffffu) > 0x7f800000u) {
        v.u = ((v.u >> 16) | 0x40u) << 16;
        return v.f;
    }
    uint32_t a = v.u & 0x7fffffffu, sg = v.u & 0x80000000u;
    if (a < 0x7f800000u) a = a + (r >> 16);
    if (a > 0x7f800000u) a = 0x7f800000u;
    v.u = (a | sg) & 0xffff0000u;
    return v.f;
}
static uint32_t dither_base(uint32_t step, uint32_t cell) { return fmix32(cell * 0x9E3779B1u ^ step * 0x85EBCA77u); }

static void codec16(const O *o, Codec16 *k) {
    const lbm_config_t *c = &o->c;
    k->on = sizeof(REAL) == sizeof(float) && (c->storage == LBM_STORE_FP16 || c->storage == LBM_STORE_BF16);
    const float ux = c->storage_u_ref[0], uy = c->storage_u_ref[1], uz = c->storage_u_ref[2];
    const float u2 = ux * ux + uy * uy + uz * uz;
    for (int q = 0; q < 19; q++) {
        float cu = 3.0f * ((float)CX[q] * ux + (float)CY[q] * uy + (float)CZ[q] * uz);
        k->sh[q] = (float)WD[q] * (1.0f + cu + 0.5f * cu * cu - 1.5f * u2);
    }
    k->scale = 1.0f;
    for (int i = 0; i < c->storage_shift; i++) k->scale *= 2.0f;
    k->inv = 1.0f / k->scale;
    k->sr = c->storage_rounding == LBM_ROUND_STOCHASTIC;
    k->half = c->storage == LBM_STORE_FP16;
}
/* the stored value of population q with value v: round((v - s_q) 2^k) read
   back as s_q + h 2^-k, with the dither word of (base, q) */
static REAL round16(const Codec16 *k, REAL v, int q, uint32_t base) {
    float d = ((float)v - k->sh[q]) * k->scale;
    uint32_t r = base + (uint32_t)q * 0x9E3779B9u;
    float h = k->half ? (k->sr ? half_sr(d, r) : half_rne(d)) : (k->sr ? bf16_sr(d, r) : bf16_rne(d));
    return (REAL)(k->sh[q] + h * k->inv);
}
static void store_round(O *o, uint32_t step) {
    Codec16 k;
    codec16(o, &k);
    if (!k.on) return;
    for (size_t id = 0; id < o->n; id++) {
        uint32_t base = dither_base(step, (uint32_t)id);
        for (int q = 0; q < 19; q++) F(q, id) = round16(&k, F(q, id), q, base);
    }
}
/* The device's fused step (air without a free surface, the box or the phi
   modes) stores populations once per step, after the collision.  Its staged
   step also stores the post-stream populations and, where they write, the
   conditioning (the values it changed), the refill (the refilled cells) and
   the initial free-surface projection (the projected links); each of those
   store points draws its dither from the step number with a salt of its
   own.  A 16-bit run on the staged path rounds at all of them. */
static int fused_path(const O *o) {
    const lbm_config_t *c = &o->c;
    if (c->kernel_path == LBM_PATH_STAGED) return 0;
    return c->physics_mode == LBM_PHYS_AIR && !c->fs_enabled && !c->include_cube && !c->phi_fraction && o->nz >= 3;
}
static int staged_rounds(const O *o) {
    Codec16 k;
    codec16(o, &k);
    return k.on && !fused_path(o);
}
static uint32_t bits_of(REAL v) {
    float f = (float)v;
    uint32_t u;
    memcpy(&u, &f, sizeof u);
    return u;
}

/* collide_KBC (CODEX:3143-3321) */
static void collide_kbc(O *o) {
    const lbm_config_t *c = &o->c;
    REAL cs2 = R(1.0) / R(3.0);
    REAL eps = R(1e-6);
    REAL om_sp = R(1.0) / R(0.95);
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                REAL rl = o->rho[id], ux = o->ux[id], uy = o->uy[id], uz = o->uz[id];
                int is_solid = LO(0, id) == 0;
                int is_gas = 0;
                if (is_water(o) && o->type[id] == LBM_FS_GAS) {
                    is_gas = 1;
                    rl = R(c->rho0);
                    ux = uy = uz = R(0);
                }
                REAL u2 = ux * ux + uy * uy + uz * uz;
                int sponge = c->enable_top_sponge == 1 && k > o->nz - 3;
                if (sponge) {
                    for (int q = 0; q < 19; q++) {
                        REAL cu = R(3.0) * ((REAL)CX[q] * ux + (REAL)CY[q] * uy + (REAL)CZ[q] * uz);
                        REAL fe = rl * o->w[q] * (R(1.0) + cu + R(0.5) * cu * cu - R(1.5) * u2);
                        REAL fv = F(q, id);
                        F(q, id) = fv - om_sp * (fv - fe);
                    }
                    continue;
                }
                REAL om = o->omega[id];
                REAL fe[19], fs[19], fp[19];
                for (int q = 0; q < 19; q++) {
                    REAL cu = R(3.0) * ((REAL)CX[q] * ux + (REAL)CY[q] * uy + (REAL)CZ[q] * uz);
                    fe[q] = rl * o->w[q] * (R(1.0) + cu + R(0.5) * cu * cu - R(1.5) * u2);
                }
                for (int q = 0; q < 19; q++) {
                    REAL fv = F(q, id);
                    REAL force_i = R(0);
                    if (is_water(o) && !is_gas) {
                        REAL force_z = -rl * local_g(o, id);
                        REAL cxi = (REAL)CX[q], cyi = (REAL)CY[q], czi = (REAL)CZ[q];
                        REAL cdu = cxi * ux + cyi * uy + czi * uz;
                        force_i = o->w[q] * (R(1.0) - R(0.5) * om) *
                                  (((czi - uz) / cs2) + (cdu * czi / (cs2 * cs2))) * force_z;
                    }
                    fs[q] = fv - om * (fv - fe[q]) + force_i;
                }
                REAL alpha = R(1.0);
                for (int q = 0; q < 19; q++)
                    if (fs[q] < eps) {
                        REAL den = fs[q] - fe[q];
                        if (den < R(0)) {
                            REAL cand = (eps - fe[q]) / den;
                            if (cand < alpha) alpha = cand;
                        }
                    }
                if (alpha < R(0)) alpha = R(0);
                if (alpha > R(1)) alpha = R(1);
                for (int q = 0; q < 19; q++) fp[q] = fe[q] + alpha * (fs[q] - fe[q]);
                if (R(c->c_backscatter) > R(0) && !is_solid) {
                    REAL Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
                    for (int q = 0; q < 19; q++) {
                        REAL fneq = fp[q] - fe[q];
                        REAL cxi = (REAL)CX[q], cyi = (REAL)CY[q], czi = (REAL)CZ[q];
                        Pxx += fneq * cxi * cxi;
                        Pyy += fneq * cyi * cyi;
                        Pzz += fneq * czi * czi;
                        Pxy += fneq * cxi * cyi;
                        Pxz += fneq * cxi * czi;
                        Pyz += fneq * cyi * czi;
                    }
                    REAL tr = (Pxx + Pyy + Pzz) / R(3.0);
                    Pxx -= tr;
                    Pyy -= tr;
                    Pzz -= tr;
                    REAL xi = R(2.0) * (REAL)bs_uniform(c->seed, (uint32_t)o->step, (uint32_t)id) - R(1.0);
                    REAL scale = R(c->c_backscatter) * xi;
                    REAL inv2 = R(1.0) / (R(2.0) * cs2 * cs2);
                    REAL delta[19];
                    for (int q = 0; q < 19; q++) {
                        REAL cxi = (REAL)CX[q], cyi = (REAL)CY[q], czi = (REAL)CZ[q];
                        REAL Qxx = cxi * cxi - cs2, Qyy = cyi * cyi - cs2, Qzz = czi * czi - cs2;
                        REAL Qxy = cxi * cyi, Qxz = cxi * czi, Qyz = cyi * czi;
                        REAL contr = (Qxx * (scale * Pxx) + Qyy * (scale * Pyy) + Qzz * (scale * Pzz) +
                                      R(2.0) * (Qxy * (scale * Pxy) + Qxz * (scale * Pxz) + Qyz * (scale * Pyz)));
                        delta[q] = o->w[q] * inv2 * contr;
                    }
                    REAL gamma = R(1.0);
                    for (int q = 0; q < 19; q++)
                        if (delta[q] < R(0)) {
                            REAL cand = (eps - fp[q]) / delta[q];
                            if (cand < gamma) gamma = cand;
                        }
                    if (gamma < R(0)) gamma = R(0);
                    if (gamma > R(1)) gamma = R(1);
                    for (int q = 0; q < 19; q++) fp[q] = fp[q] + gamma * delta[q];
                }
                for (int q = 0; q < 19; q++) F(q, id) = fp[q];
            }
}

/* collide_regularized (CODEX:3324-3416) */
static void collide_reg(O *o) {
    const lbm_config_t *c = &o->c;
    REAL cs2 = R(1.0) / R(3.0);
    REAL inv2 = R(1.0) / (R(2.0) * cs2 * cs2);
    REAL eps = R(1e-6);
#pragma omp parallel for collapse(3) schedule(static)
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                REAL rl = o->rho[id], ux = o->ux[id], uy = o->uy[id], uz = o->uz[id];
                int is_gas = 0;
                if (is_water(o) && o->type[id] == LBM_FS_GAS) {
                    is_gas = 1;
                    rl = R(c->rho0);
                    ux = uy = uz = R(0);
                }
                REAL om = o->omega[id];
                REAL u2 = ux * ux + uy * uy + uz * uz;
                REAL fe[19], fs[19];
                for (int q = 0; q < 19; q++) {
                    REAL cu = R(3.0) * ((REAL)CX[q] * ux + (REAL)CY[q] * uy + (REAL)CZ[q] * uz);
                    fe[q] = rl * o->w[q] * (R(1.0) + cu + R(0.5) * cu * cu - R(1.5) * u2);
                }
                if (is_gas) {
                    for (int q = 0; q < 19; q++) F(q, id) = fe[q];
                    continue;
                }
                REAL Pxx = 0, Pyy = 0, Pzz = 0, Pxy = 0, Pxz = 0, Pyz = 0;
                for (int q = 0; q < 19; q++) {
                    REAL fneq = F(q, id) - fe[q];
                    REAL cxi = (REAL)CX[q], cyi = (REAL)CY[q], czi = (REAL)CZ[q];
                    Pxx += fneq * cxi * cxi;
                    Pyy += fneq * cyi * cyi;
                    Pzz += fneq * czi * czi;
                    Pxy += fneq * cxi * cyi;
                    Pxz += fneq * cxi * czi;
                    Pyz += fneq * cyi * czi;
                }
                for (int q = 0; q < 19; q++) {
                    REAL cxi = (REAL)CX[q], cyi = (REAL)CY[q], czi = (REAL)CZ[q];
                    REAL Qxx = cxi * cxi - cs2, Qyy = cyi * cyi - cs2, Qzz = czi * czi - cs2;
                    REAL Qxy = cxi * cyi, Qxz = cxi * czi, Qyz = cyi * czi;
                    REAL freg = o->w[q] * inv2 *
                                (Qxx * Pxx + Qyy * Pyy + Qzz * Pzz + R(2.0) * (Qxy * Pxy + Qxz * Pxz + Qyz * Pyz));
                    REAL force_i = R(0);
                    if (is_water(o)) {
                        REAL force_z = -rl * local_g(o, id);
                        REAL cdu = cxi * ux + cyi * uy + czi * uz;
                        force_i = o->w[q] * (R(1.0) - R(0.5) * om) *
                                  (((czi - uz) / cs2) + (cdu * czi / (cs2 * cs2))) * force_z;
                    }
                    fs[q] = fe[q] + (R(1.0) - om) * freg + force_i;
                }
                REAL alpha = R(1.0);
                for (int q = 0; q < 19; q++)
                    if (fs[q] < eps) {
                        REAL den = fs[q] - fe[q];
                        if (den < R(0)) {
                            REAL cand = (eps - fe[q]) / den;
                            if (cand < alpha) alpha = cand;
                        }
                    }
                if (alpha < R(0)) alpha = R(0);
                if (alpha > R(1)) alpha = R(1);
                for (int q = 0; q < 19; q++) F(q, id) = fe[q] + alpha * (fs[q] - fe[q]);
            }
}

/* post-collision value of population q at column (si,sj) of plane sk, where
   si may lie one cell outside x for the inflow/outflow boundary (new):
   the inflow ghost holds feq(rho0, u_in), the outflow ghost copies i=nx-1. */
static REAL src_value(const O *o, int si, int sj, int sk, int q) {
    if (si < 0) {
        REAL uin = o->inflow_ux[(size_t)sk * o->ny + sj];
        return feq(o, R(o->c.rho0), uin, R(0), R(0), q);
    }
    if (si >= o->nx) si = o->nx - 1;
    return F(q, IDX(sk, sj, si));
}

/* stream (CODEX:3418-3551) */
static void stream(O *o) {
    const lbm_config_t *c = &o->c;
    int water = is_water(o);
    int xper = c->x_boundary == LBM_XB_PERIODIC;
    int xio = c->x_boundary == LBM_XB_INFLOW_OUTFLOW;
    int yper = c->y_boundary == LBM_YB_PERIODIC;
    int nx = o->nx, ny = o->ny, nz = o->nz;
    REAL rho0 = R(c->rho0);
    lbm_counters_t *cn = &o->cnt;
    int n_cut = 0, n_fs = 0, n_bridge = 0;  /* integer sums: independent of the thread split */
#pragma omp parallel for collapse(3) schedule(static) reduction(+ : n_cut, n_fs, n_bridge)
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < ny; j++)
            for (int i = 0; i < nx; i++) {
                size_t id = IDX(k, j, i);
                int dest_gas = water && o->type[id] == LBM_FS_GAS && LO(0, id) == 1;
                FN(0, id) = dest_gas ? feq(o, rho0, R(0), R(0), R(0), 0) : F(0, id);
                for (int q = 1; q < 19; q++) {
                    int sir = i - CX[q];
                    int si = sir, fluid_i = 1;
                    if (xper)
                        si = pmod(sir, nx);
                    else if (!xio) {
                        if (sir < 0) {
                            si = 0;
                            fluid_i = 0;
                        } else if (sir >= nx) {
                            si = nx - 1;
                            fluid_i = 0;
                        }
                    }
                    int sjr = j - CY[q];
                    int sj = yper ? pmod(sjr, ny) : sjr;
                    if (!yper && (sjr < 0 || sjr >= ny)) fluid_i = 0; /* solid y: bounce back */
                    int sk = k - CZ[q];
                    int qb = OPP[q];
                    if (water) {
                        int src_gas = 0;
                        if (fluid_i && sk >= 0 && sk < nz) {
                            if (si >= 0 && si < nx) {
                                size_t sid = IDX(sk, sj, si);
                                src_gas = o->type[sid] == LBM_FS_GAS && LO(0, sid) == 1;
                            }
                        } else if (sk >= nz) {
                            src_gas = 1;
                        }
                        if (dest_gas) {
                            FN(q, id) = feq(o, rho0, R(0), R(0), R(0), q);
                        } else if (!fluid_i) {
                            FN(q, id) = F(qb, id);
                            n_cut++;
                        } else if (sk >= 0 && sk < nz && !src_gas) {
                            REAL frac = LF(q, id);
                            REAL fstr = frac * src_value(o, si, sj, sk, q) + (R(1.0) - frac) * F(qb, id);
                            REAL fout = fstr;
                            if (c->fs_tracking == LBM_FST_VOF && c->vof_link_aperture) {
                                REAL la = (si >= 0 && si < nx) ? liquid_fraction(o, IDX(sk, sj, si)) : R(1);
                                if (la < R(1.0)) {
                                    REAL rfs = rho0;
                                    if (use_hbp(o) || c->fs_boundary_mode == LBM_FSB_HYDROSTATIC)
                                        rfs = fs_density_link(o, k, j, i, q);
                                    REAL fa = feq(o, rfs, o->ux[id], o->uy[id], o->uz[id], q);
                                    REAL fb = feq(o, rfs, o->ux[id], o->uy[id], o->uz[id], qb);
                                    REAL ffs = fa + fb - F(qb, id);
                                    fout = la * fstr + (R(1.0) - la) * ffs;
                                    n_fs++;
                                }
                            }
                            FN(q, id) = fout;
                            if (frac < R(1.0)) n_cut++;
                        } else if (sk < 0) {
                            FN(q, id) = F(qb, id);
                            n_cut++;
                        } else {
                            REAL rfs = rho0;
                            if (use_hbp(o) || c->fs_boundary_mode == LBM_FSB_HYDROSTATIC)
                                rfs = fs_density_link(o, k, j, i, q);
                            REAL fa = feq(o, rfs, o->ux[id], o->uy[id], o->uz[id], q);
                            REAL fb = feq(o, rfs, o->ux[id], o->uy[id], o->uz[id], qb);
                            REAL ffs = fa + fb - F(qb, id);
                            REAL fout = ffs;
                            if (c->fs_tracking == LBM_FST_VOF && c->vof_thin_gap_bridge && sk >= 0 && sk < nz &&
                                src_gas) {
                                int bir = si - CX[q];
                                int bi = bir;
                                int bj = yper ? pmod(sj - CY[q], ny) : sj - CY[q];
                                int bk = sk - CZ[q];
                                int bv = 1;
                                if (xper)
                                    bi = pmod(bir, nx);
                                else if (bir < 0 || bir >= nx)
                                    bv = 0;
                                if (!yper && (bj < 0 || bj >= ny)) bv = 0;
                                if (bk < 0 || bk >= nz) bv = 0;
                                if (bv) {
                                    size_t bid = IDX(bk, bj, bi);
                                    if (LO(0, bid) == 1 && o->type[bid] != LBM_FS_GAS) {
                                        REAL bw = o->fill[id];
                                        if (o->fill[bid] < bw) bw = o->fill[bid];
                                        bw *= R(c->vof_thin_gap_bridge_strength);
                                        if (bw > R(0)) {
                                            if (bw > R(1)) bw = R(1);
                                            fout = (R(1.0) - bw) * ffs + bw * F(q, bid);
                                            n_bridge++;
                                        }
                                    }
                                }
                            }
                            FN(q, id) = fout;
                            n_fs++;
                        }
                    } else if (!fluid_i) {
                        FN(q, id) = F(qb, id);
                        n_cut++;
                    } else if (sk >= 0 && sk < nz) {
                        REAL frac = LF(q, id);
                        FN(q, id) = frac * src_value(o, si, sj, sk, q) + (R(1.0) - frac) * F(qb, id);
                        if (frac < R(1.0)) n_cut++;
                    } else if (sk >= nz && c->top_boundary == LBM_TOP_FREE_SLIP) {
                        /* new: specular reflection at the top face */
                        FN(q, id) = src_value(o, si, sj, nz - 1, MIRZ[q]);
                    } else {
                        FN(q, id) = F(q, id);
                    }
                }
            }
    cn->boundary_cut_link_count = n_cut;
    cn->free_surface_cut_link_count = n_fs;
    cn->free_surface_thin_gap_bridge_count = n_bridge;
}

/* apply_x_solid_wall_bounceback (CODEX:3554-3580) */
static void x_wall_bounceback(O *o) {
    if (o->c.x_boundary != LBM_XB_SOLID) return;
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++) {
            size_t L = IDX(k, j, 0), Rr = IDX(k, j, o->nx - 1);
            int la = 1, ra = 1;
            if (is_water(o)) {
                la = !(o->type[L] == LBM_FS_GAS && LO(0, L) == 1);
                ra = !(o->type[Rr] == LBM_FS_GAS && LO(0, Rr) == 1);
            }
            for (int q = 1; q < 19; q++) {
                if (CX[q] > 0 && la) FN(q, L) = F(OPP[q], L);
                if (CX[q] < 0 && ra) FN(q, Rr) = F(OPP[q], Rr);
            }
        }
}

/* apply_open_boundary_conditions (CODEX:3582-3712) */
static void open_bc(O *o) {
    const lbm_config_t *c = &o->c;
    REAL rho0 = R(c->rho0);
    int nz = o->nz;
    if (c->enable_x_fringe) {
        REAL denom_i = R(4.0);
        for (int k = 0; k < nz; k++)
            for (int j = 0; j < o->ny; j++)
                for (int i = 0; i < 5 && i < o->nx; i++) {
                    size_t id = IDX(k, j, i);
                    if (!(LO(0, id) == 1 && k > nz / 2)) continue;
                    REAL s = R(1.0) - (REAL)i / denom_i;
                    REAL kr = (REAL)k / (REAL)(nz - 1);
                    REAL strength = R(0.05) * (kr * kr - R(0.25));
                    REAL sigma = strength * smoothstep01(s);
                    REAL uxt = o->u_inlet_profile[k] + R(0.0);
                    for (int q = 0; q < 19; q++) {
                        REAL ft = feq(o, rho0, uxt, R(0), R(0), q);
                        FN(q, id) = (R(1.0) - sigma) * FN(q, id) + sigma * ft;
                    }
                }
    }
    if (c->enable_top_drive) {
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(nz - 1, j, i);
                if (LO(0, id) == 1)
                    for (int q = 0; q < 19; q++) FN(q, id) = feq(o, rho0, R(c->u_top_lb), R(0), R(0), q);
            }
    }
    for (int k = 0; k < nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                if (o->near[id] != 1) continue;
                for (int q = 1; q < 19; q++) {
                    int qb = OPP[q];
                    REAL fo = LF(q, id);
                    REAL uw[3] = {R(0), R(0), R(0)};
                    if (WT(q, id) == LBM_WALL_BED) wall_velocity(o, j, i, q, uw);
                    REAL cuw = (REAL)CX[qb] * uw[0] + (REAL)CY[qb] * uw[1] + (REAL)CZ[qb] * uw[2];
                    REAL fw = F(qb, id) - R(6.0) * o->w[qb] * rho0 * cuw;
                    FN(q, id) = FN(q, id) * fo + fw * (R(1.0) - fo);
                }
            }
}

static void swap_f(O *o) {
    REAL *t = o->f;
    o->f = o->fn;
    o->fn = t;
}

/* ---- VOF pipeline (CODEX:1305-2374) ------------------------------------ */

/* compute_vof_mass_exchange (CODEX:1305-1350) */
static void vof_mass_exchange(O *o) {
    for (size_t id = 0; id < o->n; id++) {
        o->mass_delta[id] = R(0);
        o->excess[id] = R(0);
        o->excess_dir[id] = 0;
    }
    o->cnt.free_surface_vof_excess_count = 0;
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                if (!(LO(0, id) == 1 && o->type[id] == LBM_FS_INTERFACE)) continue;
                REAL dm = R(0);
                for (int q = 1; q < 19; q++) {
                    int ni, nj, nk;
                    if (!nbr(o, i, j, k, CX[q], CY[q], CZ[q], &ni, &nj, &nk)) continue;
                    size_t nid = IDX(nk, nj, ni);
                    if (LO(0, nid) == 1 && o->type[nid] != LBM_FS_GAS) {
                        REAL lw = R(1.0);
                        if (o->type[nid] == LBM_FS_INTERFACE) lw = R(0.5) * (o->fill[id] + o->fill[nid]);
                        if (lw > R(0)) dm += lw * (F(OPP[q], id) - F(q, nid));
                    }
                }
                o->mass_delta[id] = dm;
            }
}

/* apply_vof_mass_update_and_preclassify (CODEX:1353-1414) */
static void vof_preclassify(O *o) {
    const lbm_config_t *c = &o->c;
    for (size_t id = 0; id < o->n; id++) {
        int old = o->type[id];
        o->type_prev[id] = old;
        o->excess[id] = R(0);
        o->excess_dir[id] = 0;
        if (LO(0, id) == 0) {
            o->mass[id] = R(0);
            o->fill[id] = R(0);
            o->type[id] = LBM_FS_GAS;
            continue;
        }
        REAL rr = o->rho[id];
        if (rr < R(1.0e-8)) rr = R(c->rho0);
        REAL m = o->mass[id];
        if (old == LBM_FS_INTERFACE)
            m += o->mass_delta[id];
        else if (old == LBM_FS_LIQUID)
            m = rr;
        else
            m = R(0);
        REAL ex = R(0);
        int ed = 0;
        if (m < R(0)) {
            ex = m;
            ed = -1;
            m = R(0);
        } else if (m > rr) {
            ex = m - rr;
            ed = 1;
            m = rr;
        }
        REAL fill = m / rr;
        int ct = LBM_FS_INTERFACE;
        if (fill > R(1.0e-4) && fill < R(c->vof_empty_fill_threshold)) {
            ex = m;
            ed = -1;
            m = R(0);
            fill = R(0);
            ct = LBM_FS_GAS;
        } else if (fill <= R(1.0e-4)) {
            fill = R(0);
            ct = LBM_FS_GAS;
        } else if (fill >= R(1.0) - R(1.0e-4)) {
            fill = R(1.0);
            ct = LBM_FS_LIQUID;
        }
        o->mass[id] = m;
        o->excess[id] = ex;
        o->excess_dir[id] = ed;
        if ((REAL)fabs(ex) > R(1.0e-7)) o->cnt.free_surface_vof_excess_count++;
        o->fill[id] = fill;
        o->type[id] = ct;
    }
}

/* room of a receiving cell (CODEX:1487-1496) */
static REAL vof_room(const O *o, size_t nid, REAL excess, const REAL *msnap) {
    if (excess > R(0)) {
        REAL rr = o->rho[nid];
        if (rr < R(1.0e-8)) rr = R(o->c.rho0);
        return rr - msnap[nid];
    }
    return msnap[nid];
}

/* redistribute_vof_mass_excess (CODEX:1417-1574), deterministic two-phase
   form: phase A computes each donor's normal, weight and capacity sums from
   the snapshot; phase B lets each receiver add its shares in q order. */
static void vof_redistribute(O *o) {
    REAL *ms = o->s_m; /* mass snapshot */
    REAL *nxs = o->s_a, *nys = o->s_b, *nzs = o->s_c, *wsum = o->s_d, *csum = o->s_e;
    int32_t *dir = o->s_i;
    memcpy(ms, o->mass, o->n * sizeof(REAL));
    int xper = o->c.x_boundary == LBM_XB_PERIODIC, yper = o->c.y_boundary == LBM_YB_PERIODIC;
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                dir[id] = 0;
                REAL ex = o->excess[id];
                if (!((REAL)fabs(ex) > R(1.0e-12))) continue;
                int ed = o->excess_dir[id];
                if (ed == 0) ed = ex < R(0) ? -1 : 1;
                int ip = i + 1, im = i - 1;
                if (xper) {
                    ip = (i + 1) % o->nx;
                    im = (i - 1 + o->nx) % o->nx;
                } else {
                    if (ip >= o->nx) ip = o->nx - 1;
                    if (im < 0) im = 0;
                }
                int jp = j + 1, jm = j - 1;
                if (yper) {
                    jp = (j + 1) % o->ny;
                    jm = (j - 1 + o->ny) % o->ny;
                } else {
                    if (jp >= o->ny) jp = o->ny - 1;
                    if (jm < 0) jm = 0;
                }
                int kp = k + 1 >= o->nz ? o->nz - 1 : k + 1, km = k - 1 < 0 ? 0 : k - 1;
                REAL nx_ = -(o->fill[IDX(k, j, ip)] - o->fill[IDX(k, j, im)]);
                REAL ny_ = -(o->fill[IDX(k, jp, i)] - o->fill[IDX(k, jm, i)]);
                REAL nz_ = -(o->fill[IDX(kp, j, i)] - o->fill[IDX(km, j, i)]);
                REAL nm = (REAL)sqrt(nx_ * nx_ + ny_ * ny_ + nz_ * nz_);
                if (nm > R(1.0e-8)) {
                    REAL inv = R(1.0) / nm;
                    nx_ *= inv;
                    ny_ *= inv;
                    nz_ *= inv;
                }
                REAL ws = R(0), cs = R(0);
                for (int q = 1; q < 19; q++) {
                    int ni, nj, nk;
                    if (!nbr(o, i, j, k, CX[q], CY[q], CZ[q], &ni, &nj, &nk)) continue;
                    size_t nid = IDX(nk, nj, ni);
                    if (LO(0, nid) != 1) continue;
                    REAL nd = (REAL)CX[q] * nx_ + (REAL)CY[q] * ny_ + (REAL)CZ[q] * nz_;
                    int ok = ed > 0 ? o->type[nid] != LBM_FS_LIQUID : o->type[nid] != LBM_FS_GAS;
                    if (!ok) continue;
                    REAL room = vof_room(o, nid, ex, ms);
                    if (room > R(1.0e-12)) {
                        cs += room;
                        REAL wgt = ed > 0 ? nd : -nd;
                        if (wgt > R(0)) ws += wgt * room;
                    }
                }
                dir[id] = ed;
                nxs[id] = nx_;
                nys[id] = ny_;
                nzs[id] = nz_;
                wsum[id] = ws;
                csum[id] = cs;
            }
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                REAL acc = ms[id];
                if (dir[id] != 0 && !(csum[id] > R(1.0e-12))) acc += o->excess[id];
                if (LO(0, id) == 1) {
                    for (int q = 1; q < 19; q++) {
                        int si, sj, sk;
                        if (!nbr(o, i, j, k, -CX[q], -CY[q], -CZ[q], &si, &sj, &sk)) continue;
                        size_t sid = IDX(sk, sj, si);
                        if (dir[sid] == 0 || !(csum[sid] > R(1.0e-12))) continue;
                        int ed = dir[sid];
                        REAL ex = o->excess[sid];
                        int ok = ed > 0 ? o->type[id] != LBM_FS_LIQUID : o->type[id] != LBM_FS_GAS;
                        if (!ok) continue;
                        REAL room = vof_room(o, id, ex, ms);
                        if (!(room > R(1.0e-12))) continue;
                        REAL nd = (REAL)CX[q] * nxs[sid] + (REAL)CY[q] * nys[sid] + (REAL)CZ[q] * nzs[sid];
                        REAL wgt = ed > 0 ? nd : -nd;
                        REAL share = ex * room / csum[sid];
                        if (wsum[sid] > R(1.0e-12) && wgt > R(0))
                            share = ex * wgt * room / wsum[sid];
                        else if (wsum[sid] > R(1.0e-12))
                            share = R(0);
                        acc += share;
                    }
                }
                o->mass[id] = acc;
            }
    for (size_t id = 0; id < o->n; id++)
        if (dir[id] != 0) {
            o->excess[id] = R(0);
            o->excess_dir[id] = 0;
        }
}

/* finalize_vof_classification_from_mass (CODEX:1577-1624) */
static void vof_finalize(O *o) {
    const lbm_config_t *c = &o->c;
    for (size_t id = 0; id < o->n; id++) {
        o->excess[id] = R(0);
        o->excess_dir[id] = 0;
        if (LO(0, id) == 0) {
            o->mass[id] = R(0);
            o->fill[id] = R(0);
            o->type[id] = LBM_FS_GAS;
            continue;
        }
        REAL rr = o->rho[id];
        if (rr < R(1.0e-8)) rr = R(c->rho0);
        REAL m = o->mass[id];
        REAL ex = R(0);
        int ed = 0;
        if (m < R(0)) {
            ex = m;
            ed = -1;
            m = R(0);
            o->cnt.free_surface_vof_excess_count++;
        }
        if (m > rr) {
            ex = m - rr;
            ed = 1;
            m = rr;
            o->cnt.free_surface_vof_excess_count++;
        }
        REAL fill = m / rr;
        int ct = LBM_FS_INTERFACE;
        if (fill > R(1.0e-4) && fill < R(c->vof_empty_fill_threshold)) {
            ex = m;
            ed = -1;
            m = R(0);
            fill = R(0);
            ct = LBM_FS_GAS;
            o->cnt.free_surface_vof_excess_count++;
        } else if (fill <= R(1.0e-4)) {
            fill = R(0);
            ct = LBM_FS_GAS;
        } else if (fill >= R(1.0) - R(1.0e-4)) {
            fill = R(1.0);
            ct = LBM_FS_LIQUID;
        }
        o->mass[id] = m;
        o->excess[id] = ex;
        o->excess_dir[id] = ed;
        o->fill[id] = fill;
        o->type[id] = ct;
    }
}

/* enforce_vof_interface_layer (CODEX:1627-1656) */
static void vof_enforce_layer(O *o) {
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                if (!(LO(0, id) == 1 && o->type[id] == LBM_FS_LIQUID)) continue;
                int touches = 0;
                for (int q = 1; q < 19; q++) {
                    int ni, nj, nk;
                    if (!nbr(o, i, j, k, CX[q], CY[q], CZ[q], &ni, &nj, &nk)) continue;
                    size_t nid = IDX(nk, nj, ni);
                    if (LO(0, nid) == 1 && o->type[nid] == LBM_FS_GAS) touches = 1;
                }
                if (touches) {
                    o->type[id] = LBM_FS_INTERFACE;
                    o->fill[id] = R(1.0);
                    o->mass[id] = o->rho[id];
                }
            }
}

/* remove_orphan_vof_interface_cells (CODEX:1659-1776), deterministic form */
static void vof_remove_orphans(O *o) {
    const lbm_config_t *c = &o->c;
    REAL *ms = o->s_m, *fs_ = o->s_a, *capw = o->s_b, *topw = o->s_c;
    int32_t *ts = o->s_i; /* type snapshot; orphan flag stored as bit 8 */
    REAL thr = R(c->vof_orphan_empty_threshold);
    int rmax = c->vof_orphan_search_radius;
    memcpy(ms, o->mass, o->n * sizeof(REAL));
    memcpy(fs_, o->fill, o->n * sizeof(REAL));
    for (size_t id = 0; id < o->n; id++) ts[id] = o->type[id];
    o->cnt.free_surface_vof_orphan_count = 0;
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                if (!(LO(0, id) == 1 && (ts[id] & 0xff) != LBM_FS_GAS && fs_[id] > R(c->vof_empty_fill_threshold)))
                    continue;
                int wet = 0, res = 0;
                for (int q = 1; q < 19; q++) {
                    int ni, nj, nk;
                    if (!nbr(o, i, j, k, CX[q], CY[q], CZ[q], &ni, &nj, &nk)) continue;
                    size_t nid = IDX(nk, nj, ni);
                    if (LO(0, nid) == 1 && (ts[nid] & 0xff) != LBM_FS_GAS) {
                        wet++;
                        if (fs_[nid] >= thr) res++;
                    }
                }
                if (!(res == 0 && (fs_[id] < thr || wet <= c->vof_orphan_max_weak_neighbors))) continue;
                REAL m = ms[id];
                REAL cw = R(0), tw = R(0);
                for (int q = 1; q < 19; q++)
                    for (int d = 2; d <= rmax; d++) {
                        int ni, nj, nk;
                        if (!nbr(o, i, j, k, d * CX[q], d * CY[q], d * CZ[q], &ni, &nj, &nk)) continue;
                        size_t nid = IDX(nk, nj, ni);
                        if (LO(0, nid) == 1 && (ts[nid] & 0xff) != LBM_FS_GAS && fs_[nid] >= thr) {
                            REAL rr = o->rho[nid];
                            if (rr < R(1.0e-8)) rr = R(c->rho0);
                            REAL room = rr - ms[nid];
                            int di = d * CX[q], dj = d * CY[q], dk = d * CZ[q];
                            REAL d2 = (REAL)(di * di + dj * dj + dk * dk);
                            tw += fs_[nid] / d2;
                            if (room > R(1.0e-12)) cw += room / d2;
                        }
                    }
                if (m > R(1.0e-12) && (cw > R(1.0e-12) || tw > R(1.0e-12))) {
                    ts[id] |= 0x100;
                    capw[id] = cw;
                    topw[id] = tw;
                    o->cnt.free_surface_vof_orphan_count++;
                }
            }
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                int is_orphan = (ts[id] & 0x100) != 0;
                REAL acc = is_orphan ? R(0) : ms[id];
                if (LO(0, id) == 1 && (ts[id] & 0xff) != LBM_FS_GAS && fs_[id] >= thr) {
                    REAL rr = o->rho[id];
                    if (rr < R(1.0e-8)) rr = R(c->rho0);
                    REAL room = rr - ms[id];
                    for (int q = 1; q < 19; q++)
                        for (int d = 2; d <= rmax; d++) {
                            int si, sj, sk;
                            if (!nbr(o, i, j, k, -d * CX[q], -d * CY[q], -d * CZ[q], &si, &sj, &sk)) continue;
                            size_t sid = IDX(sk, sj, si);
                            if (!(ts[sid] & 0x100)) continue;
                            int di = d * CX[q], dj = d * CY[q], dk = d * CZ[q];
                            REAL d2 = (REAL)(di * di + dj * dj + dk * dk);
                            REAL wgt = R(0), den = capw[sid];
                            if (capw[sid] > R(1.0e-12)) {
                                if (room > R(1.0e-12)) wgt = room / d2;
                            } else {
                                wgt = fs_[id] / d2;
                                den = topw[sid];
                            }
                            if (den > R(1.0e-12) && wgt > R(0)) acc += ms[sid] * wgt / den;
                        }
                }
                o->mass[id] = acc;
                if (is_orphan) {
                    o->fill[id] = R(0);
                    o->type[id] = LBM_FS_GAS;
                }
            }
}

/* vof_unresolved_detached_cell (CODEX:1779-1810) on the snapshot arrays */
static int vof_detached(const O *o, int k, int j, int i, const int32_t *ts, const REAL *fs_) {
    size_t id = IDX(k, j, i);
    if (!(LO(0, id) == 1 && ts[id] != LBM_FS_GAS && fs_[id] > R(1.0e-4))) return 0;
    int wet = 0, res = 0;
    for (int q = 1; q < 19; q++) {
        int ni, nj, nk;
        if (!nbr(o, i, j, k, CX[q], CY[q], CZ[q], &ni, &nj, &nk)) continue;
        size_t nid = IDX(nk, nj, ni);
        if (LO(0, nid) == 1 && ts[nid] != LBM_FS_GAS) {
            wet++;
            if (fs_[nid] >= R(o->c.vof_orphan_empty_threshold)) res++;
        }
    }
    return wet <= o->c.vof_detached_max_wet_neighbors && res <= o->c.vof_detached_max_resolved_neighbors;
}

static const int FD[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};

/* vof_gas_neighbor_rate (CODEX:1813-1831) */
static REAL gas_rate(const O *o, int k, int j, int i, int d, REAL rate, const int32_t *ts) {
    if (!(rate > R(0))) return R(0);
    int ni, nj, nk;
    if (!nbr(o, i, j, k, FD[d][0], FD[d][1], FD[d][2], &ni, &nj, &nk)) return R(0);
    size_t nid = IDX(nk, nj, ni);
    if (LO(0, nid) == 1 && ts[nid] == LBM_FS_GAS) return rate;
    return R(0);
}

/* mass moved by cell (k,j,i) in face direction d (CODEX:1834-1857 with the
   callers CODEX:1861-1914 / 1917-1993); mode 0 = advection, 1 = compaction */
static REAL detached_move(const O *o, int k, int j, int i, int d, int mode, const int32_t *ts, const REAL *fs_,
                          const REAL *ms) {
    size_t id = IDX(k, j, i);
    if (!vof_detached(o, k, j, i, ts, fs_)) return R(0);
    REAL m = ms[id];
    REAL ux = o->ux[id], uy = o->uy[id], uz = o->uz[id];
    REAL sgn[6] = {ux, -ux, uy, -uy, uz, -uz};
    if (mode == 0) {
        if (!(m > R(1.0e-12))) return R(0);
        REAL r[6], rs = R(0);
        for (int e = 0; e < 6; e++) r[e] = gas_rate(o, k, j, i, e, sgn[e], ts);
        rs = r[0] + r[1] + r[2] + r[3] + r[4] + r[5];
        if (!(rs > R(1.0e-12))) return R(0);
        REAL scale = R(1.0);
        if (rs > R(0.95)) scale = R(0.95) / rs;
        if (!(r[d] > R(0))) return R(0);
        REAL mv = m * r[d] * scale;
        return mv > R(1.0e-12) ? mv : R(0);
    }
    REAL rr = o->rho[id];
    if (rr < R(1.0e-8)) rr = R(o->c.rho0);
    REAL fill = m / rr;
    if (!(fill > R(1.0e-4) && fill < R(o->c.vof_detached_residual_fill_threshold))) return R(0);
    int best = -1;
    REAL br = R(0);
    for (int e = 0; e < 6; e++) {
        REAL r = gas_rate(o, k, j, i, e, sgn[e], ts);
        if (r > br) {
            br = r;
            best = e;
        }
    }
    if (!(br > R(1.0e-12)) || best != d) return R(0);
    REAL mv = m * R(1.0) * R(1.0);
    return mv > R(1.0e-12) ? mv : R(0);
}

static void vof_detached_pass(O *o, int mode) {
    const lbm_config_t *c = &o->c;
    if (mode == 0) {
        o->cnt.free_surface_vof_detached_advect_count = 0;
        o->cnt.free_surface_vof_detached_advect_mass = 0.0;
    } else {
        o->cnt.free_surface_vof_detached_compress_count = 0;
        o->cnt.free_surface_vof_detached_compress_mass = 0.0;
    }
    if (!c->vof_detached_advection) return;
    if (mode == 1 && !(c->vof_detached_residual_fill_threshold > 0.0f)) return;
    REAL *ms = o->s_m, *fs_ = o->s_a;
    int32_t *ts = o->s_i;
    memcpy(ms, o->mass, o->n * sizeof(REAL));
    memcpy(fs_, o->fill, o->n * sizeof(REAL));
    for (size_t id = 0; id < o->n; id++) ts[id] = o->type[id];
    REAL *mv = o->s_mv;
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                for (int d = 0; d < 6; d++) mv[6 * id + d] = detached_move(o, k, j, i, d, mode, ts, fs_, ms);
            }
    int32_t cntv = 0;
    double mass_sum = 0.0;
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                REAL out = R(0), moved = R(0);
                for (int d = 0; d < 6; d++) {
                    out -= mv[6 * id + d];
                    moved += mv[6 * id + d];
                }
                REAL in = R(0);
                for (int d = 0; d < 6; d++) {
                    int si, sj, sk;
                    if (!nbr(o, i, j, k, -FD[d][0], -FD[d][1], -FD[d][2], &si, &sj, &sk)) continue;
                    in += mv[6 * IDX(sk, sj, si) + d];
                }
                REAL dm = out + in;
                if ((REAL)fabs(dm) > R(1.0e-12)) o->mass[id] = ms[id] + dm;
                if (moved > R(1.0e-12)) {
                    cntv++;
                    mass_sum += (double)moved;
                }
            }
    if (mode == 0) {
        o->cnt.free_surface_vof_detached_advect_count = cntv;
        o->cnt.free_surface_vof_detached_advect_mass = mass_sum;
    } else {
        o->cnt.free_surface_vof_detached_compress_count = cntv;
        o->cnt.free_surface_vof_detached_compress_mass = mass_sum;
    }
}

/* vof_gas_fraction_at (CODEX:1996-2009) */
static REAL gas_fraction(const O *o, size_t id) {
    REAL g = R(0);
    if (LO(0, id) == 1) {
        if (o->type[id] == LBM_FS_GAS)
            g = R(1.0);
        else if (o->fill[id] < R(1.0)) {
            g = R(1.0) - o->fill[id];
            if (g < R(0)) g = R(0);
            if (g > R(1)) g = R(1);
        }
    }
    return g;
}

/* collapse_trapped_gas_if_due (CODEX:2011-2143): flood fill to fixed point,
   FP64 sums, apply. */
static void vof_collapse(O *o) {
    const lbm_config_t *c = &o->c;
    lbm_counters_t *cn = &o->cnt;
    cn->free_surface_vof_collapse_cell_count = 0;
    cn->free_surface_vof_collapse_candidate_volume_m3 = 0.0;
    cn->free_surface_vof_collapse_applied_volume_m3 = 0.0;
    cn->free_surface_vof_collapse_removal_scale = 0.0;
    cn->free_surface_vof_collapse_candidate_mass = 0.0;
    cn->free_surface_vof_collapse_exterior_capacity = 0.0;
    if (!c->vof_collapse_trapped_gas || !is_vof(o) || (o->step % c->vof_collapse_interval) != 0) return;
    int32_t *mark = o->ext_gas;
    /* breadth-first reachability from the seeded top layer */
    int32_t *queue = o->s_i;
    size_t head = 0, tail = 0;
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                int m = gas_fraction(o, id) > R(1.0e-6) && k == o->nz - 1;
                mark[id] = m;
                if (m) queue[tail++] = (int32_t)id;
            }
    while (head < tail) {
        size_t id = (size_t)queue[head++];
        int i = (int)(id % (size_t)o->nx), j = (int)((id / (size_t)o->nx) % (size_t)o->ny),
            k = (int)(id / ((size_t)o->nx * (size_t)o->ny));
        for (int q = 1; q < 7; q++) {
            int ni, nj, nk;
            if (!nbr(o, i, j, k, CX[q], CY[q], CZ[q], &ni, &nj, &nk)) continue;
            size_t nid = IDX(nk, nj, ni);
            if (mark[nid] == 0 && gas_fraction(o, nid) > R(1.0e-6)) {
                mark[nid] = 1;
                queue[tail++] = (int32_t)nid;
            }
        }
    }
    double cmass = 0.0, cvol = 0.0, cap = 0.0;
    REAL dx = R(c->dx);
    for (size_t id = 0; id < o->n; id++) {
        REAL g = gas_fraction(o, id);
        if (!(g > R(1.0e-6))) continue;
        if (mark[id] == 0) {
            REAL rr = o->rho[id];
            if (rr < R(1.0e-8)) rr = R(c->rho0);
            cmass += (double)(g * rr);
            cvol += (double)(g * dx * dx * dx);
        } else if (o->type[id] != LBM_FS_GAS) {
            REAL m = o->mass[id];
            if (m > R(0)) cap += (double)m;
        }
    }
    cn->free_surface_vof_collapse_candidate_mass = cmass;
    cn->free_surface_vof_collapse_candidate_volume_m3 = cvol;
    cn->free_surface_vof_collapse_exterior_capacity = cap;
    if (cvol > 0.0 && cvol <= (double)c->vof_collapse_max_volume_m3 && cap >= cmass && cap > 1.0e-12) {
        cn->free_surface_vof_collapse_applied_volume_m3 = cvol;
        cn->free_surface_vof_collapse_removal_scale = cmass / cap;
    }
    REAL scale = (REAL)cn->free_surface_vof_collapse_removal_scale;
    if (!(scale > R(0))) return;
    for (size_t id = 0; id < o->n; id++) {
        REAL g = gas_fraction(o, id);
        if (!(g > R(1.0e-6))) continue;
        REAL rr = o->rho[id];
        if (rr < R(1.0e-8)) rr = R(c->rho0);
        if (mark[id] == 0) {
            o->mass[id] = rr;
            o->fill[id] = R(1.0);
            o->type[id] = LBM_FS_LIQUID;
            cn->free_surface_vof_collapse_cell_count++;
        } else if (o->type[id] != LBM_FS_GAS) {
            REAL m = o->mass[id] * (R(1.0) - scale);
            if (m < R(0)) m = R(0);
            REAL fill = m / rr;
            int ct = LBM_FS_INTERFACE;
            if (fill <= R(1.0e-4)) {
                fill = R(0);
                m = R(0);
                ct = LBM_FS_GAS;
            } else if (fill >= R(1.0) - R(1.0e-4)) {
                fill = R(1.0);
                m = rr;
                ct = LBM_FS_LIQUID;
            }
            o->mass[id] = m;
            o->fill[id] = fill;
            o->type[id] = ct;
        }
    }
}

/* update_free_surface_height_from_vof (CODEX:2146-2161) */
static void vof_height(O *o) {
    REAL dx = R(o->c.dx);
    for (int j = 0; j < o->ny; j++)
        for (int i = 0; i < o->nx; i++) {
            REAL depth = R(0);
            for (int k = 0; k < o->nz; k++) {
                size_t id = IDX(k, j, i);
                if (LO(0, id) == 1) depth += o->fill[id] * LF(0, id) * dx;
            }
            o->fs_h_last[ID2(j, i)] = o->fs_h[ID2(j, i)];
            o->fs_h[ID2(j, i)] = o->water_h[ID2(j, i)] + depth;
        }
}

/* ---- height-function free-surface tracking (CODEX:1209-1232, 2376-2429) -- */

/* classify_free_surface_cells: fill = clamp((h - k dx) / dx, 0, 1) of the
   column height, types by the 1e-4 thresholds, closed cells gas */
static void fs_classify(O *o) {
    const lbm_config_t *c = &o->c;
    const REAL dx = R(c->dx);
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                int old = o->type[id];
                REAL z_bottom = (REAL)k * dx;
                REAL fill = (o->fs_h[ID2(j, i)] - z_bottom) / dx;
                if (fill < R(0.0)) fill = R(0.0);
                if (fill > R(1.0)) fill = R(1.0);
                int t = LBM_FS_INTERFACE;
                if (fill <= R(1e-4))
                    t = LBM_FS_GAS;
                else if (fill >= R(1.0) - R(1e-4))
                    t = LBM_FS_LIQUID;
                if (LO(0, id) == 0) {
                    fill = R(0.0);
                    t = LBM_FS_GAS;
                }
                o->fill[id] = fill;
                o->type_prev[id] = old;
                o->type[id] = t;
            }
}

/* compute_free_surface_column_flux: q = sum_k fill u vel_scale dx over the
   open, non-gas cells of a column */
static void fs_column_flux(O *o) {
    const lbm_config_t *c = &o->c;
    const REAL vs = R(c->vel_scale), dx = R(c->dx);
    for (int j = 0; j < o->ny; j++)
        for (int i = 0; i < o->nx; i++) {
            REAL qx = R(0), qy = R(0);
            for (int k = 0; k < o->nz; k++) {
                size_t id = IDX(k, j, i);
                if (LO(0, id) == 1 && o->type[id] != LBM_FS_GAS) {
                    REAL fl = o->fill[id];
                    qx += fl * o->ux[id] * vs * dx;
                    qy += fl * o->uy[id] * vs * dx;
                }
            }
            o->fs_qx[ID2(j, i)] = qx;
            o->fs_qy[ID2(j, i)] = qy;
        }
}

/* divergence of a column flux along one axis at index a of n: central
   differences on a periodic axis, half-sum face fluxes with zero flux
   through the walls of a solid one */
static REAL flux_div(const REAL *q, size_t stride, int a, int n, int solid, REAL dx) {
    if (solid) {
        REAL fl = R(0), fr = R(0);
        if (a > 0) fl = R(0.5) * (q[(size_t)(a - 1) * stride] + q[(size_t)a * stride]);
        if (a < n - 1) fr = R(0.5) * (q[(size_t)a * stride] + q[(size_t)(a + 1) * stride]);
        return (fr - fl) / dx;
    }
    int ap = (a + 1) % n, am = (a - 1 + n) % n;
    return (q[(size_t)ap * stride] - q[(size_t)am * stride]) * (R(0.5) / dx);
}

/* advect_free_surface_height: h += dt (-(dq_x/dx + dq_y/dy)), kept between
   half a cell above the bed and half a cell below the lid (height_kinematic);
   the solitary wave at time t (prescribed_solitary); height_static keeps h.
   y walls (new) take the x-wall form. */
/* solitary_surface_height_at_x (CODEX:1166-1182) in float: the KdV wave
   A sech^2(kappa (x - x0 - dir c t)), kappa = sqrt(3 A / (4 h0^3)), the offset
   folded into the periodic domain; exp is det_exp */
static REAL solitary_height(const O *o, REAL xp, REAL t) {
    const lbm_config_t *c = &o->c;
    const REAL lx = (REAL)o->nx * R(c->dx);
    const REAL amp = R(c->solitary[0]), h0 = R(c->solitary[1]);
    const REAL kappa = (REAL)sqrt((double)(R(3.0) * amp / (R(4.0) * h0 * h0 * h0)));
    const REAL center = R(c->solitary[2]) + R(c->solitary[3]) * R(c->solitary[4]) * t;
    REAL xrel = xp - center;
    if (c->x_boundary == LBM_XB_PERIODIC) {
        if (xrel > R(0.5) * lx) xrel -= lx;
        if (xrel < R(-0.5) * lx) xrel += lx;
    }
    REAL arg = kappa * xrel;
    REAL ep = (REAL)det_exp_d((double)arg);
    REAL em = (REAL)det_exp_d((double)(-arg));
    REAL sech = R(2.0) / (ep + em);
    return R(c->free_surface_base_m) + amp * sech * sech;
}

static void fs_advect_height(O *o, REAL t) {
    const lbm_config_t *c = &o->c;
    const REAL dx = R(c->dx), dt = R(c->dt_phys);
    const REAL lz = (REAL)o->nz * dx;
    for (int j = 0; j < o->ny; j++)
        for (int i = 0; i < o->nx; i++) {
            size_t cid = ID2(j, i);
            REAL h_old = o->fs_h[cid], h_new = h_old;
            if (c->fs_mode == LBM_FSM_PRESCRIBED_SOLITARY) {
                h_new = solitary_height(o, ((REAL)i + R(0.5)) * dx, t);
            } else if (c->fs_mode == LBM_FSM_HEIGHT_KINEMATIC) {
                REAL dq_dx = flux_div(o->fs_qx + (size_t)j * o->nx, 1, i, o->nx, c->x_boundary == LBM_XB_SOLID, dx);
                REAL dq_dy = flux_div(o->fs_qy + i, (size_t)o->nx, j, o->ny, c->y_boundary == LBM_YB_SOLID, dx);
                REAL dh_dt = -(dq_dx + dq_dy);
                h_new = h_old + dt * dh_dt;
                REAL min_h = o->water_h[cid] + R(0.5) * dx;
                REAL max_h = lz - R(0.5) * dx;
                if (h_new < min_h) h_new = min_h;
                if (h_new > max_h) h_new = max_h;
            }
            o->fs_h_last[cid] = h_old;
            o->fs_h[cid] = h_new;
        }
}

/* refill_new_free_surface_cells (CODEX:2163-2316) */
static void vof_refill(O *o) {
    const lbm_config_t *c = &o->c;
    Codec16 cd;
    codec16(o, &cd);
    const int rnd = staged_rounds(o);  /* the refilled cells are stored (rounded) */
    o->cnt.free_surface_refill_count = 0;
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                if (!(o->type_prev[id] == LBM_FS_GAS && o->type[id] != LBM_FS_GAS && LO(0, id) == 1)) continue;
                REAL ws = 0, fws = 0, ra = 0, xa = 0, ya = 0, za = 0, rf = 0, xf = 0, yf = 0, zf = 0;
                REAL fna[19], fnf[19];
                for (int q = 0; q < 19; q++) fna[q] = fnf[q] = R(0);
                for (int q = 1; q < 19; q++) {
                    int ni, nj, nk;
                    if (!nbr(o, i, j, k, CX[q], CY[q], CZ[q], &ni, &nj, &nk)) continue;
                    size_t nid = IDX(nk, nj, ni);
                    if (!(LO(0, nid) == 1 && o->type_prev[nid] != LBM_FS_GAS && o->type[nid] != LBM_FS_GAS)) continue;
                    REAL wn = o->w[q], fl = o->fill[nid];
                    REAL rn = o->rho[nid], uxn = o->ux[nid], uyn = o->uy[nid], uzn = o->uz[nid];
                    fws += wn;
                    rf += wn * rn;
                    xf += wn * uxn;
                    yf += wn * uyn;
                    zf += wn * uzn;
                    REAL lw = wn * fl;
                    ws += lw;
                    ra += lw * rn;
                    xa += lw * uxn;
                    ya += lw * uyn;
                    za += lw * uzn;
                    if (c->refill_mode == LBM_REFILL_NONEQ)
                        for (int qq = 0; qq < 19; qq++) {
                            REAL fneq = F(qq, nid) - feq(o, rn, uxn, uyn, uzn, qq);
                            fna[qq] += lw * fneq;
                            fnf[qq] += wn * fneq;
                        }
                }
                /* VOF: density from the liquid neighbours; height tracking:
                   the hydrostatic density of the column (CODEX:2225-2246) */
                const int vof = is_vof(o);
                REAL rr = vof ? R(c->rho0) : fs_density_cell(o, k, j, i);
                REAL urx = R(c->u_current_lb), ury = R(0), urz = R(0);
                if (ws > R(0)) {
                    REAL inv = R(1.0) / ws;
                    if (vof) rr = ra * inv;
                    urx = xa * inv;
                    ury = ya * inv;
                    urz = za * inv;
                } else if (fws > R(0)) {
                    REAL inv = R(1.0) / fws;
                    if (vof) rr = rf * inv;
                    urx = xf * inv;
                    ury = yf * inv;
                    urz = zf * inv;
                } else if (vof) {
                    rr = R(c->rho0);
                }
                REAL fv[19];
                REAL rp = R(0);
                for (int q = 0; q < 19; q++) {
                    REAL fr = feq(o, rr, urx, ury, urz, q);
                    if (c->refill_mode == LBM_REFILL_NONEQ) {
                        if (ws > R(0))
                            fr += fna[q] / ws;
                        else if (fws > R(0))
                            fr += fnf[q] / fws;
                    } else if (c->refill_mode == LBM_REFILL_DIRECTIONAL) {
                        int si, sj, sk;
                        if (nbr(o, i, j, k, -CX[q], -CY[q], -CZ[q], &si, &sj, &sk)) {
                            size_t sid = IDX(sk, sj, si);
                            if (LO(0, sid) == 1 && o->type_prev[sid] != LBM_FS_GAS && o->type[sid] != LBM_FS_GAS)
                                fr = F(q, sid);
                        }
                    }
                    fv[q] = fr;
                    rp += fr;
                }
                if (c->refill_mode == LBM_REFILL_DIRECTIONAL) {
                    REAL scale = R(1.0);
                    if (rp > R(1.0e-8)) scale = rr / rp;
                    REAL r2 = 0, mx = 0, my = 0, mz = 0;
                    for (int q = 0; q < 19; q++) {
                        REAL fvv = fv[q] * scale;
                        F(q, id) = fvv;
                        r2 += fvv;
                        mx += fvv * (REAL)CX[q];
                        my += fvv * (REAL)CY[q];
                        mz += fvv * (REAL)CZ[q];
                    }
                    o->rho[id] = r2;
                    if (r2 > R(1.0e-8)) {
                        o->ux[id] = mx / r2;
                        o->uy[id] = my / r2;
                        o->uz[id] = mz / r2;
                    } else {
                        o->ux[id] = urx;
                        o->uy[id] = ury;
                        o->uz[id] = urz;
                    }
                } else {
                    o->rho[id] = rr;
                    o->ux[id] = urx;
                    o->uy[id] = ury;
                    o->uz[id] = urz;
                    for (int q = 0; q < 19; q++) F(q, id) = fv[q];
                }
                if (rnd) {
                    const uint32_t base = dither_base((uint32_t)o->step ^ SALT_REFILL, (uint32_t)id);
                    for (int q = 0; q < 19; q++) F(q, id) = round16(&cd, F(q, id), q, base);
                }
                o->cnt.free_surface_refill_count++;
            }
}

/* repair_vof_coalescence_links (CODEX:2319-2373) */
static void vof_repair(O *o) {
    o->cnt.free_surface_coalescence_repair_count = 0;
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                if (!(LO(0, id) == 1 && o->type[id] != LBM_FS_GAS && o->type_prev[id] != LBM_FS_GAS)) continue;
                int rep = 0;
                for (int q = 1; q < 19; q++) {
                    int si, sj, sk;
                    if (!nbr(o, i, j, k, -CX[q], -CY[q], -CZ[q], &si, &sj, &sk)) continue;
                    size_t sid = IDX(sk, sj, si);
                    if (LO(0, sid) == 1 && o->type_prev[sid] == LBM_FS_GAS && o->type[sid] != LBM_FS_GAS) {
                        F(q, id) = F(q, sid);
                        rep = 1;
                        o->cnt.free_surface_coalescence_repair_count++;
                    }
                }
                if (rep) {
                    REAL r = 0, mx = 0, my = 0, mz = 0;
                    for (int q = 0; q < 19; q++) {
                        REAL fq = F(q, id);
                        r += fq;
                        mx += fq * (REAL)CX[q];
                        my += fq * (REAL)CY[q];
                        mz += fq * (REAL)CZ[q];
                    }
                    o->rho[id] = r;
                    if (r > R(1.0e-8)) {
                        o->ux[id] = mx / r;
                        o->uy[id] = my / r;
                        o->uz[id] = mz / r;
                    }
                }
            }
}

/* initialize_vof_mass_from_fill (CODEX:1289-1302) */
static void vof_init_mass(O *o) {
    for (size_t id = 0; id < o->n; id++) {
        if (LO(0, id) == 1 && o->type[id] != LBM_FS_GAS) {
            o->mass[id] = o->fill[id] * o->rho[id];
        } else {
            o->mass[id] = R(0);
            o->fill[id] = R(0);
            o->type[id] = LBM_FS_GAS;
        }
        o->type_prev[id] = o->type[id];
        o->mass_delta[id] = R(0);
        o->excess[id] = R(0);
        o->excess_dir[id] = 0;
    }
}

/* ---- exported API ------------------------------------------------------- */

void *API(orc_create)(const lbm_config_t *cfg) {
    if (!cfg || cfg->version != LBM_CONFIG_VERSION || cfg->nx < 2 || cfg->ny < 2 || cfg->nz < 3) return NULL;
    O *o = (O *)calloc(1, sizeof(O));
    o->c = *cfg;
    o->nx = cfg->nx;
    o->ny = cfg->ny;
    o->nz = cfg->nz;
    o->n = (size_t)o->nx * o->ny * o->nz;
    o->n2 = (size_t)o->nx * o->ny;
    for (int q = 0; q < 19; q++) {
        o->w[q] = (REAL)WD[q];
        MIRZ[q] = q;
        for (int p = 0; p < 19; p++)
            if (CX[p] == CX[q] && CY[p] == CY[q] && CZ[p] == -CZ[q]) MIRZ[q] = p;
    }
    size_t n = o->n, n2 = o->n2;
#define AR(p, cnt) p = (REAL *)calloc((cnt), sizeof(REAL))
#define AI(p, cnt) p = (int32_t *)calloc((cnt), sizeof(int32_t))
    AR(o->f, 19 * n);
    AR(o->fn, 19 * n);
    AR(o->rho, n);
    AR(o->ux, n);
    AR(o->uy, n);
    AR(o->uz, n);
    AR(o->omega, n);
    AR(o->nu_t, n);
    AR(o->smag, n);
    AR(o->phi, n);
    AI(o->lo, 19 * n);
    AI(o->wt, 19 * n);
    AR(o->lf, 19 * n);
    AI(o->near, n);
    AI(o->obst, n);
    AI(o->obst_prev, n);
    AR(o->water_h, n2);
    AR(o->water_h_last, n2);
    AR(o->u_wave, n2);
    AR(o->v_wave, n2);
    AR(o->w_wave, n2);
    AR(o->piecewise_bed_z, o->nx);
    AR(o->u_inlet_profile, o->nz);
    AR(o->inflow_ux, (size_t)o->nz * o->ny);
    AR(o->fill, n);
    AR(o->mass, n);
    AR(o->mass_delta, n);
    AR(o->excess, n);
    AR(o->det_delta, n);
    AR(o->pref_head, n);
    AR(o->grav_frac, n);
    AI(o->type, n);
    AI(o->type_prev, n);
    AI(o->excess_dir, n);
    AI(o->ext_gas, n);
    AR(o->fs_h, n2);
    AR(o->fs_h_last, n2);
    AR(o->fs_qx, n2);
    AR(o->fs_qy, n2);
    AR(o->s_m, n);
    AR(o->s_a, n);
    AR(o->s_b, n);
    AR(o->s_c, n);
    AR(o->s_d, n);
    AR(o->s_e, n);
    AR(o->s_mv, 6 * n);
    AI(o->s_i, n);
    /* init_constants (CODEX:991-1080) */
    REAL om0 = R(1.0) / R(cfg->tau0);
    for (size_t id = 0; id < n; id++) {
        o->rho[id] = R(cfg->rho0);
        o->fill[id] = R(1.0);
        o->type[id] = LBM_FS_LIQUID;
        o->type_prev[id] = LBM_FS_LIQUID;
        o->mass[id] = R(cfg->rho0);
        o->omega[id] = om0;
    }
    for (size_t id = 0; id < n2; id++) {
        int i = (int)(id % (size_t)o->nx);
        REAL h0 = R(cfg->offset_m);
        if (cfg->bed_profile == LBM_BED_PIECEWISE) h0 = o->piecewise_bed_z[i];
        o->water_h[id] = h0;
        o->water_h_last[id] = h0;
        o->fs_h[id] = R(cfg->free_surface_base_m);
        o->fs_h_last[id] = R(cfg->free_surface_base_m);
    }
    return o;
}

void API(orc_destroy)(void *p) {
    O *o = (O *)p;
    if (!o) return;
    void *ptrs[] = {o->f,    o->fn,       o->rho,      o->ux,       o->uy,        o->uz,       o->omega,
                    o->nu_t, o->smag,     o->phi,      o->lo,       o->wt,        o->lf,       o->near,
                    o->obst, o->obst_prev, o->water_h, o->water_h_last, o->u_wave, o->v_wave, o->w_wave,
                    o->piecewise_bed_z, o->u_inlet_profile, o->inflow_ux, o->fill, o->mass, o->mass_delta,
                    o->excess, o->det_delta, o->pref_head, o->grav_frac, o->type, o->type_prev, o->excess_dir,
                    o->ext_gas, o->fs_h, o->fs_h_last, o->fs_qx, o->fs_qy, o->s_m, o->s_a, o->s_b, o->s_c, o->s_d, o->s_e, o->s_mv, o->s_i};
    for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(ptrs[i]);
    free(o);
}

/* field table: pointer, element count, is_int */
static void *field_ptr(O *o, int field, size_t *count, int *is_int) {
    size_t n = o->n, n2 = o->n2;
    *is_int = 0;
    switch (field) {
        case LBM_FIELD_F: *count = 19 * n; return o->f;
        case LBM_FIELD_RHO: *count = n; return o->rho;
        case LBM_FIELD_UX: *count = n; return o->ux;
        case LBM_FIELD_UY: *count = n; return o->uy;
        case LBM_FIELD_UZ: *count = n; return o->uz;
        case LBM_FIELD_OMEGA: *count = n; return o->omega;
        case LBM_FIELD_PHI: *count = n; return o->phi;
        case LBM_FIELD_LATTICE_OPEN: *count = 19 * n; *is_int = 1; return o->lo;
        case LBM_FIELD_LATTICE_OPEN_FRAC: *count = 19 * n; return o->lf;
        case LBM_FIELD_WALL_TYPE: *count = 19 * n; *is_int = 1; return o->wt;
        case LBM_FIELD_NEAR_OBSTACLE: *count = n; *is_int = 1; return o->near;
        case LBM_FIELD_OBSTACLE: *count = n; *is_int = 1; return o->obst;
        case LBM_FIELD_WATER_H: *count = n2; return o->water_h;
        case LBM_FIELD_U_WAVE: *count = n2; return o->u_wave;
        case LBM_FIELD_V_WAVE: *count = n2; return o->v_wave;
        case LBM_FIELD_W_WAVE: *count = n2; return o->w_wave;
        case LBM_FIELD_FS_FILL: *count = n; return o->fill;
        case LBM_FIELD_FS_TYPE: *count = n; *is_int = 1; return o->type;
        case LBM_FIELD_FS_TYPE_PREV: *count = n; *is_int = 1; return o->type_prev;
        case LBM_FIELD_FS_MASS: *count = n; return o->mass;
        case LBM_FIELD_FS_MASS_EXCESS: *count = n; return o->excess;
        case LBM_FIELD_FS_H: *count = n2; return o->fs_h;
        case LBM_FIELD_PRESSURE_REF_HEAD: *count = n; return o->pref_head;
        case LBM_FIELD_GRAVITY_FRACTION: *count = n; return o->grav_frac;
        case LBM_FIELD_U_INLET_PROFILE: *count = (size_t)o->nz; return o->u_inlet_profile;
        case LBM_FIELD_INFLOW_UX: *count = (size_t)o->nz * o->ny; return o->inflow_ux;
        case LBM_FIELD_PIECEWISE_BED_Z: *count = (size_t)o->nx; return o->piecewise_bed_z;
        case LBM_FIELD_NU_T: *count = n; return o->nu_t;
        case LBM_FIELD_SMAG: *count = n; return o->smag;
        case LBM_FIELD_VOF_EXTERIOR_GAS: *count = n; *is_int = 1; return o->ext_gas;
        default: return NULL;
    }
}

/* copy a field in (dir=0) or out (dir=1); host data is REAL or int32 */
int API(orc_field)(void *p, int field, void *data, size_t bytes, int dir) {
    O *o = (O *)p;
    size_t count;
    int is_int;
    void *ptr = field_ptr(o, field, &count, &is_int);
    if (!ptr) return LBM_E_ARG;
    size_t need = count * (is_int ? sizeof(int32_t) : sizeof(REAL));
    if (bytes != need) return LBM_E_ARG;
    if (dir == 0) {
        memcpy(ptr, data, need);
        if (field == LBM_FIELD_F) store_round(o, (uint32_t)o->step);
        if (field == LBM_FIELD_PIECEWISE_BED_Z && o->c.bed_profile == LBM_BED_PIECEWISE)
            for (size_t id = 0; id < o->n2; id++) {
                o->water_h[id] = o->piecewise_bed_z[id % (size_t)o->nx];
                o->water_h_last[id] = o->water_h[id];
            }
    } else
        memcpy(data, ptr, need);
    return 0;
}

/* geometry at time t (CODEX:3779-3781) */
void API(orc_geometry)(void *p, float t) {
    O *o = (O *)p;
    bed_update(o, (REAL)t);
    compute_phi(o);
    build_lattice_open(o);
    for (size_t id = 0; id < o->n; id++) o->obst[id] = LO(0, id) == 0 ? 1 : 0;
}

/* init_fields tail + populations (CODEX:2712, 2789-2818, 2829-2875), then the
   VOF initialisation of main() (CODEX:3789-3792). */
/* initialize_free_surface_boundary_populations (CODEX:2829-2875): in every
   open wet cell, each population streamed from an open gas cell (or from
   above the lid) is replaced by the gas-pressure closure of the pull stream,
   feq_q(rho_fs, u) + feq_qbar(rho_fs, u) - f_qbar; sources outside solid
   x (or y) walls are not gas */
static void fs_init_projection(O *o) {
    const lbm_config_t *c = &o->c;
    Codec16 cd;
    codec16(o, &cd);
    const int rnd = staged_rounds(o);
    memcpy(o->fn, o->f, 19 * o->n * sizeof(REAL));
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                if (!(LO(0, id) == 1 && o->type[id] != LBM_FS_GAS)) continue;
                for (int q = 1; q < 19; q++) {
                    int si = i - CX[q], sj = j - CY[q], sk = k - CZ[q];
                    int fluid = 1;
                    if (c->x_boundary == LBM_XB_PERIODIC)
                        si = (si + o->nx) % o->nx;
                    else if (si < 0 || si >= o->nx)
                        fluid = 0;
                    if (c->y_boundary == LBM_YB_PERIODIC)
                        sj = (sj + o->ny) % o->ny;
                    else if (sj < 0 || sj >= o->ny)
                        fluid = 0;
                    int gas = 0;
                    if (!fluid)
                        gas = 0;
                    else if (sk >= o->nz)
                        gas = 1;
                    else if (sk >= 0) {
                        size_t sid = IDX(sk, sj, si);
                        gas = o->type[sid] == LBM_FS_GAS && LO(0, sid) == 1;
                    }
                    if (!gas) continue;
                    const int qb = OPP[q];
                    REAL rho_fs = R(c->rho0);
                    if (use_hbp(o) || c->fs_boundary_mode == LBM_FSB_HYDROSTATIC) rho_fs = fs_density_link(o, k, j, i, q);
                    REAL ux = o->ux[id], uy = o->uy[id], uz = o->uz[id];
                    REAL fp = feq(o, rho_fs, ux, uy, uz, q) + feq(o, rho_fs, ux, uy, uz, qb) - F(qb, id);
                    if (rnd) fp = round16(&cd, fp, q, dither_base(SALT_INIT, (uint32_t)id));  /* a projected link is stored */
                    o->fn[(size_t)q * o->n + id] = fp;
                }
            }
    memcpy(o->f, o->fn, 19 * o->n * sizeof(REAL));
}

void API(orc_init_populations)(void *p) {
    O *o = (O *)p;
    const lbm_config_t *c = &o->c;
    for (int k = 0; k < o->nz; k++)
        for (int j = 0; j < o->ny; j++)
            for (int i = 0; i < o->nx; i++) {
                size_t id = IDX(k, j, i);
                o->obst[id] = LO(0, id) == 0 ? 1 : 0;
                if (LO(0, id) == 0) {
                    o->ux[id] = o->uy[id] = o->uz[id] = R(0);
                    o->rho[id] = R(c->rho0);
                    if (WT(0, id) == LBM_WALL_BED) {
                        o->ux[id] = o->u_wave[ID2(j, i)];
                        o->uy[id] = o->v_wave[ID2(j, i)];
                        o->uz[id] = o->w_wave[ID2(j, i)];
                    }
                }
            }
    if (is_water(o)) update_pressure_reference(o);
    for (size_t id = 0; id < o->n; id++) {
        REAL ux = o->ux[id], uy = o->uy[id], uz = o->uz[id];
        if (is_water(o) && LO(0, id) == 1 && o->type[id] != LBM_FS_GAS) uz += R(0.5) * local_g(o, id);
        REAL u2 = ux * ux + uy * uy + uz * uz;
        for (int q = 0; q < 19; q++) {
            REAL cu = R(3.0) * ((REAL)CX[q] * ux + (REAL)CY[q] * uy + (REAL)CZ[q] * uz);
            F(q, id) = o->rho[id] * o->w[q] * (R(1.0) + cu + R(0.5) * cu * cu - R(1.5) * u2);
        }
    }
    store_round(o, 0);
    if (is_water(o) && c->init_fs_populations) fs_init_projection(o);
    memcpy(o->fn, o->f, 19 * o->n * sizeof(REAL));
    if (is_vof(o)) {
        vof_init_mass(o);
        vof_enforce_layer(o);
        vof_height(o);
    }
    o->step = 0;
}

/* one step of the main loop (CODEX:3873-3925) */
static void one_step(O *o) {
    const lbm_config_t *c = &o->c;
    int64_t step = o->step + 1;
    o->step = step;
    float t = (float)((double)(step + 1) * c->dt_phys);
    bed_update(o, (REAL)t);
    compute_phi(o);
    build_lattice_open(o);
    if (is_water(o) && !is_vof(o)) fs_classify(o);
    if (is_water(o)) update_pressure_reference(o);
    bed_populations(o);
    macro(o);
    if (c->use_les) les(o);
    if (c->collision == LBM_COLL_REGULARIZED)
        collide_reg(o);
    else
        collide_kbc(o);
    store_round(o, (uint32_t)o->step);
    stream(o);
    x_wall_bounceback(o);
    open_bc(o);
    swap_f(o);
    if (staged_rounds(o)) store_round(o, (uint32_t)o->step ^ SALT_STREAM);
    if (is_water(o)) {
        macro(o);
        if (is_vof(o)) {
            vof_mass_exchange(o);
            vof_preclassify(o);
            for (int r = 0; r < 3; r++) {
                vof_redistribute(o);
                vof_finalize(o);
            }
            vof_enforce_layer(o);
            vof_remove_orphans(o);
            vof_detached_pass(o, 0);
            vof_finalize(o);
            vof_detached_pass(o, 1);
            vof_finalize(o);
            vof_redistribute(o);
            vof_finalize(o);
            vof_enforce_layer(o);
            vof_collapse(o);
            vof_enforce_layer(o);
            vof_height(o);
            vof_refill(o);
            vof_repair(o);
        } else {
            fs_column_flux(o);
            fs_advect_height(o, (REAL)t);
            fs_classify(o);
            vof_refill(o);
        }
    }
}

void API(orc_step)(void *p, int nsteps) {
    O *o = (O *)p;
    for (int s = 0; s < nsteps; s++) one_step(o);
}

void API(orc_get_counters)(void *p, lbm_counters_t *out) {
    O *o = (O *)p;
    *out = o->cnt;
}

int64_t API(orc_get_step)(void *p) { return ((O *)p)->step; }

/* FP64 reductions matching lbm_reduce */
double API(orc_reduce)(void *p, int what) {
    O *o = (O *)p;
    double s = 0.0;
    if (what == LBM_REDUCE_MASS) {
        for (size_t id = 0; id < o->n; id++)
            if (LO(0, id) == 1)
                for (int q = 0; q < 19; q++) s += (double)F(q, id);
    } else if (what == LBM_REDUCE_VOF_MASS) {
        for (size_t id = 0; id < o->n; id++) s += (double)o->mass[id];
    } else if (what == LBM_REDUCE_VOF_VOLUME) {
        for (size_t id = 0; id < o->n; id++) s += (double)o->fill[id];
    }
    return s;
}

/* Threads of the per-cell loops, for the CPU-baseline timing (the results do
   not depend on it: every parallel loop writes only its own cell, and its
   counters are integer sums). */
void API(orc_set_threads)(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}
int API(orc_get_threads)(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Run one reference kernel by itself (unit tests pin each against closed
   forms): 0 bed update at time t, 1 phi, 2 lattice open, 3 pressure
   reference, 4 bed populations, 5 macro, 6 LES, 7 collision, 8 stream,
   9 x-wall bounce-back, 10 open boundary conditions, 11 swap f/f_new,
   20 VOF mass exchange, 21 preclassify, 22 redistribute, 23 finalize,
   24 interface layer, 25 orphans, 26 detached advection, 27 detached
   compaction, 28 trapped-gas collapse, 29 column height, 30 refill,
   31 coalescence repair, 32 height-function classification, 33 column
   flux, 34 height advection. */
int API(orc_stage)(void *p, int stage, float t) {
    O *o = (O *)p;
    switch (stage) {
        case 0: bed_update(o, (REAL)t); break;
        case 1: compute_phi(o); break;
        case 2: build_lattice_open(o); break;
        case 3: update_pressure_reference(o); break;
        case 4: bed_populations(o); break;
        case 5: macro(o); break;
        case 6: les(o); break;
        case 7:
            if (o->c.collision == LBM_COLL_REGULARIZED)
                collide_reg(o);
            else
                collide_kbc(o);
            break;
        case 8: stream(o); break;
        case 9: x_wall_bounceback(o); break;
        case 10: open_bc(o); break;
        case 11: swap_f(o); break;
        case 20: vof_mass_exchange(o); break;
        case 21: vof_preclassify(o); break;
        case 22: vof_redistribute(o); break;
        case 23: vof_finalize(o); break;
        case 24: vof_enforce_layer(o); break;
        case 25: vof_remove_orphans(o); break;
        case 26: vof_detached_pass(o, 0); break;
        case 27: vof_detached_pass(o, 1); break;
        case 28: vof_collapse(o); break;
        case 29: vof_height(o); break;
        case 30: vof_refill(o); break;
        case 31: vof_repair(o); break;
        case 32: fs_classify(o); break;
        case 33: fs_column_flux(o); break;
        case 34: fs_advect_height(o, (REAL)t); break;
        default: return LBM_E_ARG;
    }
    return 0;
}

/* populations after streaming of the current state (f_new) for tests */
int API(orc_field_fnew)(void *p, void *data, size_t bytes) {
    O *o = (O *)p;
    if (bytes != 19 * o->n * sizeof(REAL)) return LBM_E_ARG;
    memcpy(data, o->fn, bytes);
    return 0;
}

void API(orc_set_step)(void *p, int64_t step) { ((O *)p)->step = step; }

/* the 16-bit roundings of the storage codec on raw values (half: binary16,
   else bfloat16; sr: stochastic with the words r, else nearest-even); the
   results are returned as floats (exactly representable) */
void API(orc_round16)(int half, int sr, const float *x, const uint32_t *r, float *out, size_t n) {
    for (size_t i = 0; i < n; i++)
        out[i] = half ? (sr ? half_sr(x[i], r[i]) : half_rne(x[i])) : (sr ? bf16_sr(x[i], r[i]) : bf16_rne(x[i]));
}
