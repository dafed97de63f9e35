/*
 * lbm.h -- C ABI of the MI355X-native D3Q19 LBM time step (liblbm.so).
 *
 * The reference solver (LBM_3D_LES_air_over_waves_CODEX.py) has no function
 * API: its hot path is a fixed sequence of Taichi kernels on module-global
 * fields, driven by main() (CODEX:3757-3959).  Each entry point below replaces
 * one part of that surface; the comment on each names the reference code it
 * stands in for.  All pointers are plain host pointers; no device pointer
 * escapes the library.  Every function returns 0 on success and a negative
 * LBM_E_* code on failure; lbm_last_error() then describes the failure.
 *
 * Threading: one handle owns one HIP stream and is not thread-safe; distinct
 * handles are independent.
 */
#ifndef LBM_H
#define LBM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_CONFIG_VERSION 8

/* ---- error codes ---------------------------------------------------- */
#define LBM_OK 0
#define LBM_E_ARG (-1)      /* bad argument (size, field id, null)       */
#define LBM_E_CONFIG (-2)   /* config rejected (version, dims, modes)     */
#define LBM_E_HIP (-3)      /* HIP runtime error                          */
#define LBM_E_NOMEM (-4)    /* device allocation failed                   */
#define LBM_E_STATE (-5)    /* call not valid in the current state        */
#define LBM_E_UNSUPPORTED (-6)

/* ---- enumerations (values mirror the reference's option strings) ---- */
/* physics_mode: LBM_PHYSICS_MODE (CODEX:48) */
#define LBM_PHYS_AIR 0
#define LBM_PHYS_WATER 1
/* collision: LBM_COLLISION (CODEX:491) */
#define LBM_COLL_KBC 0
#define LBM_COLL_REGULARIZED 1
/* x_boundary: LBM_X_BOUNDARY (CODEX:200) + inflow_outflow (new) */
#define LBM_XB_PERIODIC 0
#define LBM_XB_SOLID 1
#define LBM_XB_INFLOW_OUTFLOW 2
/* y_boundary: periodic as in the reference; solid is new */
#define LBM_YB_PERIODIC 0
#define LBM_YB_SOLID 1
/* top_boundary: "keep" is the reference's out-of-domain rule (CODEX:3549-3551);
   free_slip (specular reflection) is new */
#define LBM_TOP_KEEP 0
#define LBM_TOP_FREE_SLIP 1
/* bed_profile: LBM_BED_PROFILE (CODEX:205) + moving_sine1 (single harmonic, new) */
#define LBM_BED_MOVING_SINE 0
#define LBM_BED_FLAT 1
#define LBM_BED_PIECEWISE 2
#define LBM_BED_MOVING_SINE1 3
/* storage of the populations in HBM (new) */
#define LBM_STORE_FP32 0
#define LBM_STORE_FP16 1 /* (f_i - s_i) * 2^shift as IEEE half */
#define LBM_STORE_BF16 2 /* (f_i - s_i) * 2^shift as bfloat16   */
#define LBM_STORE_FP64 3 /* FP64 storage and FP64 arithmetic    */
/* Reduced storage: s_i = feq_i(1, storage_u_ref) (= w_i when the reference
 * velocity is zero), rounded into the 16-bit format either to nearest even
 * or stochastically with a counter-based dither of (cell, step, i). */
#define LBM_ROUND_NEAREST 0
#define LBM_ROUND_STOCHASTIC 1
/* kernel path */
#define LBM_PATH_AUTO 0
#define LBM_PATH_STAGED 1 /* one kernel per reference stage */
#define LBM_PATH_FUSED 2  /* single fused z-marching pass (air mode) */
/* free_surface_mode: LBM_FREE_SURFACE_MODE (CODEX:52) */
#define LBM_FSM_NONE 0
#define LBM_FSM_HEIGHT_STATIC 1
#define LBM_FSM_HEIGHT_KINEMATIC 2
#define LBM_FSM_PRESCRIBED_SOLITARY 3
/* free_surface_tracking: LBM_FREE_SURFACE_TRACKING (CODEX:87) */
#define LBM_FST_HEIGHT 0
#define LBM_FST_VOF 1
/* pressure_formulation: LBM_PRESSURE_FORMULATION (CODEX:79) */
#define LBM_PF_TOTAL 0
#define LBM_PF_HYDROSTATIC_BALANCED 1
#define LBM_PF_VOF_COMPONENT_BALANCED 2
/* free_surface_boundary: LBM_FREE_SURFACE_BOUNDARY (CODEX:93) */
#define LBM_FSB_CELL 0
#define LBM_FSB_HYDROSTATIC 1
/* refill: LBM_FREE_SURFACE_REFILL (CODEX:99) */
#define LBM_REFILL_EQUILIBRIUM 0
#define LBM_REFILL_NONEQ 1
#define LBM_REFILL_DIRECTIONAL 2

/* wall types (CODEX:441-445) and free-surface cell types (CODEX:447-449) */
#define LBM_WALL_NONE 0
#define LBM_WALL_BED 1
#define LBM_WALL_CUBE 2
#define LBM_WALL_XMIN 3
#define LBM_WALL_XMAX 4
#define LBM_WALL_YMIN 5 /* new: solid y walls */
#define LBM_WALL_YMAX 6
#define LBM_FS_GAS 0
#define LBM_FS_LIQUID 1
#define LBM_FS_INTERFACE 2

/*
 * Run configuration: the derived constants of CODEX:35-529 in the precision
 * the reference's kernels see them (Taichi f32), plus the options that are new
 * in this build.  Built by the Python config layer from LBM_* variables.
 */
typedef struct lbm_config {
    int32_t version; /* = LBM_CONFIG_VERSION */
    int32_t nx, ny, nz;

    int32_t physics_mode, collision, use_les;
    int32_t x_boundary, y_boundary, top_boundary;
    int32_t bed_profile, include_bed, wall_velocity_link;
    int32_t enable_top_drive, enable_x_fringe, enable_top_sponge;
    int32_t storage, storage_shift, storage_rounding, kernel_path;

    /* water / free surface */
    int32_t fs_enabled, fs_tracking, pressure_formulation, fs_boundary_mode;
    int32_t refill_mode, vof_link_aperture, vof_thin_gap_bridge;
    int32_t vof_detached_advection, vof_collapse_trapped_gas;
    int32_t vof_orphan_search_radius, vof_orphan_max_weak_neighbors;
    int32_t vof_detached_max_wet_neighbors, vof_detached_max_resolved_neighbors;
    int32_t vof_collapse_interval, vof_collapse_flood_sweeps;
    int32_t fs_mode; /* LBM_FSM_*: how the height-function surface moves */
    int32_t include_cube; /* LBM_OBSTACLE_MODE bed_cube / cube (CODEX:372-400) */
    int32_t phi_fraction, phi_flags; /* LBM_BOUNDARY_GEOMETRY phi_fraction / phi (CODEX:194-198) */
    int32_t init_fs_populations;     /* LBM_INIT_FREE_SURFACE_POPULATIONS (CODEX:2829-2875) */
    uint32_t seed; /* backscatter counter-based RNG seed */

    float dx, rho0, tau0, nu0;
    float cs_les, delta_les, nu_t_max, kappa, z0_phys;
    float u_top_lb, u_current_lb, g_lb, g_body_lb;
    float c_backscatter;
    float offset_m, bed_amp_m, bed_wavelength_m, wave_speed, wave_period;
    float depth_swl_m, vel_scale, bed_phase_speed_mps;
    float free_surface_base_m;
    float vof_empty_fill_threshold, vof_orphan_empty_threshold;
    float vof_detached_residual_fill_threshold, vof_thin_gap_bridge_strength;
    float vof_collapse_max_volume_m3;
    float storage_u_ref[3]; /* reference velocity of the FP16/BF16 shift */
    float cube_half[3];     /* half sizes (m) of the box obstacle */
    float cube_center[3];   /* centre (m) */
    float cube_rot[4];      /* cos / sin of yaw, cos / sin of pitch */
    float solitary[5];      /* amplitude, depth h0, x0 (m), direction (+-1), celerity (m/s) (CODEX:290-325) */
    float pad0;
    double dt_phys; /* per-step time, the host forms t=(step+1)*dt_phys */
} lbm_config_t;

/* ---- field identifiers for upload / download ---------------------------
 * Shapes are C order with x fastest: 3-D (nz,ny,nx); per-link (19,nz,ny,nx);
 * 2-D (ny,nx).  "R" fields are float32 (float64 when storage==FP64),
 * "I" fields are int32.                                                     */
#define LBM_FIELD_F 0                /* R (19,nz,ny,nx) populations after streaming */
#define LBM_FIELD_RHO 1              /* R */
#define LBM_FIELD_UX 2               /* R */
#define LBM_FIELD_UY 3               /* R */
#define LBM_FIELD_UZ 4               /* R */
#define LBM_FIELD_OMEGA 5            /* R omegaLoc */
#define LBM_FIELD_PHI 6              /* R */
#define LBM_FIELD_LATTICE_OPEN 7     /* I (19,...) */
#define LBM_FIELD_LATTICE_OPEN_FRAC 8 /* R (19,...) */
#define LBM_FIELD_WALL_TYPE 9        /* I (19,...) */
#define LBM_FIELD_NEAR_OBSTACLE 10   /* I */
#define LBM_FIELD_OBSTACLE 11        /* I */
#define LBM_FIELD_WATER_H 12         /* R 2-D */
#define LBM_FIELD_U_WAVE 13          /* R 2-D */
#define LBM_FIELD_V_WAVE 14          /* R 2-D */
#define LBM_FIELD_W_WAVE 15          /* R 2-D */
#define LBM_FIELD_FS_FILL 16         /* R */
#define LBM_FIELD_FS_TYPE 17         /* I */
#define LBM_FIELD_FS_TYPE_PREV 18    /* I */
#define LBM_FIELD_FS_MASS 19         /* R */
#define LBM_FIELD_FS_MASS_EXCESS 20  /* R */
#define LBM_FIELD_FS_H 21            /* R 2-D */
#define LBM_FIELD_PRESSURE_REF_HEAD 22 /* R */
#define LBM_FIELD_GRAVITY_FRACTION 23 /* R */
#define LBM_FIELD_U_INLET_PROFILE 24 /* R (nz) */
#define LBM_FIELD_INFLOW_UX 25       /* R (nz,ny) inflow plane (new) */
#define LBM_FIELD_PIECEWISE_BED_Z 26 /* R (nx) */
#define LBM_FIELD_NU_T 27            /* R */
#define LBM_FIELD_SMAG 28            /* R */
#define LBM_FIELD_VOF_EXTERIOR_GAS 29 /* I */
#define LBM_FIELD_COUNT 30

/* ---- counters: the reference's shape-() fields (CODEX:600-617) --------- */
typedef struct lbm_counters {
    int32_t boundary_cut_link_count;
    int32_t boundary_refill_count;
    int32_t free_surface_refill_count;
    int32_t free_surface_coalescence_repair_count;
    int32_t free_surface_thin_gap_bridge_count;
    int32_t free_surface_cut_link_count;
    int32_t free_surface_vof_excess_count;
    int32_t free_surface_vof_orphan_count;
    int32_t free_surface_vof_collapse_cell_count;
    int32_t free_surface_vof_detached_advect_count;
    int32_t free_surface_vof_detached_compress_count;
    int32_t pad0;
    double free_surface_vof_collapse_candidate_volume_m3;
    double free_surface_vof_collapse_applied_volume_m3;
    double free_surface_vof_collapse_removal_scale;
    double free_surface_vof_collapse_candidate_mass;
    double free_surface_vof_collapse_exterior_capacity;
    double free_surface_vof_detached_advect_mass;
    double free_surface_vof_detached_compress_mass;
} lbm_counters_t;

/* reductions (FP64 on the device) */
#define LBM_REDUCE_MASS 0        /* sum of populations over open cells      */
#define LBM_REDUCE_VOF_MASS 1    /* sum of free_surface_mass                */
#define LBM_REDUCE_VOF_VOLUME 2  /* sum of free_surface_fill (cells)        */
#define LBM_REDUCE_MAX_SPEED 3   /* max |u| over active cells (CODEX:3930-3940) */
#define LBM_REDUCE_NONFINITE 4   /* count of non-finite populations          */

/* ---- parameters of the host half of the init sequence ------------------
 * (CODEX:3777-3792): the options of the initial state that the step itself
 * does not need.  Filled by the Python config layer (lbm_init.init_params)
 * or by a C caller. */
#define LBM_INIT_PARAMS_VERSION 1
#define LBM_INIT_REFERENCE 0     /* init_fields: inlet profile x (1 + TI (2U - 1)) (CODEX:2714-2749) */
#define LBM_INIT_UNIFORM_NOISE 1 /* new: u0 + U(-a, a) per component */
#define LBM_INIT_LOGLAW_NOISE 2  /* new: log-law mean x (1 + n U(-1, 1)), also the inflow plane */
#define LBM_FSI_FLAT 0           /* LBM_FREE_SURFACE_INITIAL (CODEX:1197-1207) */
#define LBM_FSI_GAUSSIAN 1
#define LBM_FSI_SOLITARY 2
#define LBM_SOLVEL_NONE 0        /* LBM_SOLITARY_INITIAL_VELOCITY (CODEX:2756-2800) */
#define LBM_SOLVEL_SHALLOW 1
#define LBM_SOLVEL_SHALLOW_DIVFREE 2
typedef struct lbm_init_params {
    int32_t version; /* = LBM_INIT_PARAMS_VERSION */
    int32_t initial_state, fs_initial, vof_block, linear_wave_pressure, solitary_velocity;
    int32_t hydrostatic_balanced, component_balanced; /* use_*_balanced_pressure (CODEX:79-86) */
    int32_t pad0;
    uint64_t seed; /* numpy.random.default_rng(seed) draws */
    double dx, offset_m, free_surface_base_m, free_surface_depth_m, rho0, g_lb, u_current_lb, u_top_lb, tau0, kappa;
    double vel_scale, turbulence_intensity;
    double init_u_lu, init_noise_lu, inflow_u_lu, inflow_z0_cells, inflow_noise;
    double gaussian_amp_m, gaussian_sigma_m, gaussian_center_x_m;
    double solitary_amp_m, solitary_x0_m, solitary_direction, solitary_celerity_phys, solitary_kappa_m_inv;
    double vof_block_size_m[3], vof_block_center_m[3];
} lbm_init_params_t;

typedef struct lbm_handle lbm_handle_t;

/* Module import + field allocation (CODEX:35-620): validate the config and
   allocate every device buffer of one case. */
int lbm_create(const lbm_config_t *cfg, lbm_handle_t **out);

/* Release the case (Taichi frees fields at process exit; no counterpart). */
void lbm_destroy(lbm_handle_t *h);

/* Geometry at time t: update_wave_bed_and_velocities(t),
   compute_phi_slope_corrected(), build_lattice_open_from_free_surface_periodic()
   (CODEX:1082-1162, 2432-2490, 2492-2686).  Used by initialisation. */
int lbm_geometry(lbm_handle_t *h, float t);

/* init_fields() populations (CODEX:2808-2818) from the uploaded rho/ux/uy/uz:
   f = feq(rho, u + 0.5 g) with the Guo shift on wet cells; f_new = f
   (initialize_free_surface_boundary_populations + copy_post_and_swap,
   CODEX:2829-2875, 3714-3727).  Water/VOF: also
   initialize_vof_mass_from_fill + enforce_vof_interface_layer +
   update_free_surface_height_from_vof (CODEX:1290-1303, 1627-1656, 2146-2161). */
int lbm_init_populations(lbm_handle_t *h);

/* The whole init sequence of main() (CODEX:3777-3792) in the library, for
   callers without Python: geometry at t = 0, the host initial fields
   (lbm_initial_fields) from the t = 0 solid mask and bed, their upload, and
   lbm_init_populations.  With a piecewise-linear bed the caller uploads
   LBM_FIELD_PIECEWISE_BED_Z first, as the reference reads it at import. */
int lbm_initialize(lbm_handle_t *h, const lbm_init_params_t *p);

/* Host half of the init sequence (init_fields CODEX:2714-2800, the free
   surface CODEX:1197-1287, prepare_initial_density_field CODEX:743-818)
   into caller buffers, without a device: obstacle0 (nz,ny,nx) the t = 0
   solid mask, bed (ny,nx) the t = 0 bed height (R) or NULL for the flat
   level; outputs R (float, or double with FP64 storage): rho, ux, uy, uz
   (nz,ny,nx), u_inlet (nz), inflow (nz,ny); water runs also fill (R), type
   (int32) and fs_h (ny,nx, R), else they may be NULL. */
int lbm_initial_fields(const lbm_config_t *cfg, const lbm_init_params_t *p, const int32_t *obstacle0, const void *bed, void *rho,
                       void *ux, void *uy, void *uz, void *u_inlet, void *inflow, void *fill, int32_t *type, void *fs_h);

/* nsteps iterations of the step loop body (CODEX:3873-3925); the first step
   advanced is number step_index+1 of the run. */
int lbm_step(lbm_handle_t *h, int32_t nsteps);

/* Several independent cases advanced together (no counterpart in the
   reference: the ensemble server's merged launch grid).  Handle hs[i]
   advances nsteps[i] steps, exactly as lbm_step(hs[i], nsteps[i]) would
   (bit for bit); each step of all handles that still have steps is one
   launch whose grid holds the tiles of every case.  The handles must be
   distinct, on the fused path, and share the storage type and the
   collision model (LBM_E_ARG otherwise).  The work runs on hs[0]'s stream,
   ordered after every handle's earlier work and before its later work. */
int lbm_step_batch(lbm_handle_t *const *hs, const int32_t *nsteps, int32_t n);

/* Host <-> device copies of one field: `.from_numpy()` / `.to_numpy()`. */
int lbm_upload_field(lbm_handle_t *h, int32_t field, const void *src, size_t bytes);
int lbm_download_field(lbm_handle_t *h, int32_t field, void *dst, size_t bytes);

/* `counter[None]` reads of the shape-() fields.  As in the reference, each
   counter describes the most recent step (the last step of the latest
   lbm_step call); the stream's link counters are counted on that step only. */
int lbm_get_counters(lbm_handle_t *h, lbm_counters_t *out);

/* FP64 device reductions (the health print CODEX:3930-3955 and the mass /
   volume diagnostics of run_ab_case.py:286-679).  Deterministic: fixed
   groups of cells and a fixed summation tree (partial sums in a buffer of
   the handle); a post-collision state is streamed on the fly. */
int lbm_reduce(lbm_handle_t *h, int32_t what, double *out);

/* lbm_reduce that allocates nothing (no counterpart in the reference): for
   callers that cannot allocate device memory (code under HIP graph capture
   of its own, tight serving loops, harnesses that forbid hipMalloc).  The
   partial sums go to a static device array of the library instead of the
   handle's buffer, so grids up to LBM_REDUCE_NOALLOC_MAX_CELLS cells are
   accepted (larger ones return LBM_E_ARG).  The array is one per GPU and
   held until the process ends; calls that overlap on one GPU (from several
   threads, or on handles whose streams run concurrently) race on it and
   must be serialised by the caller.  Same results as lbm_reduce, bit for
   bit. */
#define LBM_REDUCE_NOALLOC_MAX_CELLS ((size_t)1 << 26)
int lbm_reduce_noalloc(lbm_handle_t *h, int32_t what, double *out);

/* Number of steps advanced so far. */
int lbm_get_step(lbm_handle_t *h, int64_t *out);

/* Block until the handle's stream is idle (Taichi syncs on .to_numpy()). */
int lbm_sync(lbm_handle_t *h);

/* Description of the last failure on this handle (or of the last failed
   lbm_create when h is NULL). */
const char *lbm_last_error(lbm_handle_t *h);

/* Average duration (ms) of the dominant step kernel over the last lbm_step
   call, measured with HIP events on the handle's stream (no counterpart). */
int lbm_last_step_timing(lbm_handle_t *h, double *ms_per_step, double *ms_main_kernel);

/* Bytes of device memory the case holds (admission control; no counterpart). */
int lbm_device_bytes(const lbm_config_t *cfg, size_t *out);

/* Tuning and instrumentation knobs (no counterpart in the reference):
   "timing" (0/1: record HIP events around every step kernel),
   "block_x" (64, 128 or 256: x extent of a block of the fused passes;
   default from nx),
   "graph" (0/1: on the fused path, an lbm_step call of several steps is
   captured into a HIP graph and launched as one graph; the executable
   graph is kept and updated by the next call; ignored with "timing"),
   "wavefront", "wave_lag", "xcd_strips" (the fused step's wavefront
   schedule), "resident" (0: the wavefront schedule, 1: the tile-resident
   grid, 2, the default: the faster of the two, timed on the first
   lbm_step call of at least 18 steps, which then waits for the device),
   "vof_band" (0/1, default 1: the VOF passes of the staged step skip the
   rows of the grid the step's own change marks show settled; exact, the
   results are bit-identical with 0),
   "flood_sync" (0/1, default 1: the trapped-gas flood fill is stopped at
   its fixed point by the host, which reads its flag every 16 sweeps; 0
   launches every sweep and never waits).
   The wavefront and tile-resident schedules let a block wait for other
   blocks of its launch; while several handles exist in the process, those
   launches are chained so that one runs at a time.  Handles with "graph"
   and separate processes stepping the same GPU at the same time should use
   "wavefront" 0 and "resident" 0. */
int lbm_set_option(lbm_handle_t *h, const char *key, int64_t value);

/* Current value of an lbm_set_option knob ("resident": 2 while the choice
   by timing is pending, then the schedule chosen; no counterpart). */
int lbm_get_option(lbm_handle_t *h, const char *key, int64_t *out);

/* Resolved execution path of the handle: LBM_PATH_STAGED or LBM_PATH_FUSED. */
int lbm_get_path(lbm_handle_t *h, int32_t *out);

/* *idle = 1 when all work queued on the handle's stream has finished, 0
   otherwise; never blocks (no counterpart: used by the ensemble server). */
int lbm_query(lbm_handle_t *h, int32_t *idle);

/* ---- ensemble server (no counterpart in the reference) ------------------
 * A native stepping runtime for many independent cases on one GPU.  Worker
 * threads own the cases handed to them and advance each on its own stream in
 * chunks of chunk_steps, so that the device work of all active cases
 * overlaps; a case is reported finished once its last step has completed on
 * the device.  Between lbm_server_add and the poll that reports it finished,
 * the handle belongs to the server: the caller must not use it.  Admission
 * (by device footprint, lbm_device_bytes), initialisation and result
 * download stay with the caller (the Python EnsembleServer). */
typedef struct lbm_server lbm_server_t;

/* Start n_workers (1..64) stepping threads. */
int lbm_server_create(int32_t chunk_steps, int32_t n_workers, lbm_server_t **out);

/* Batching (before the first lbm_server_add): the fused-path cases go to one
   more thread that advances every group of them sharing a storage type and
   collision model with lbm_step_batch, one launch per step for the whole
   group (a merged launch grid); the other cases stay with the workers.
   on = 0 after batching started is LBM_E_STATE. */
int lbm_server_batch(lbm_server_t *s, int32_t on);

/* Hand over an initialised case to advance by nsteps; ticket is the caller's
   identifier, reported back by lbm_server_poll.  LBM_E_STATE if the handle
   is already held by the server (a job on it has not finished).  Sets the
   handle's options "wavefront", "resident" and "flood_sync" to 0 (cases
   stepped on concurrent streams: no waiting launches, no host waits). */
int lbm_server_add(lbm_server_t *s, lbm_handle_t *h, int32_t nsteps, int64_t ticket);

/* Move up to max finished tickets into tickets[] (*n_done of them) together
   with their status (0, or the error code of a failed lbm_step). */
int lbm_server_poll(lbm_server_t *s, int64_t *tickets, int32_t *status, int32_t max, int32_t *n_done);

/* Cases still advancing. */
int lbm_server_active(lbm_server_t *s, int32_t *n_active);

/* Stop the workers after the queued chunks of every case have completed
   (cases not yet finished are abandoned, their handles stay valid) and free
   the server. */
int lbm_server_destroy(lbm_server_t *s);

#ifdef __cplusplus
}
#endif

#endif /* LBM_H */
