// lbm_internal.h -- interface between the C ABI (lbm_capi.hip) and the
// kernel launchers (lbm_kernels.hip).  Device buffers are passed as untyped
// pointers; the launchers cast them to the arithmetic type of the run
// (double when storage == LBM_STORE_FP64, float otherwise).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/lbm.h"

namespace lbm {

// integer counters kept on the device (indices into DevBufs::icnt)
enum {
    IC_BOUNDARY_CUT = 0,
    IC_FS_CUT,
    IC_FS_BRIDGE,
    IC_BOUNDARY_REFILL,
    IC_FS_REFILL,
    IC_FS_REPAIR,
    IC_VOF_EXCESS,
    IC_VOF_ORPHAN,
    IC_VOF_COLLAPSE_CELLS,
    IC_VOF_ADVECT,
    IC_VOF_COMPRESS,
    IC_FLOOD_CHANGED,      // set by a flood-fill sweep that marked a cell
    IC_FLOOD_UNCONVERGED,  // sweeps exhausted before the fixed point
    IC_REDIST_DONORS,      // nonzero when a redistribution round has a donor (two slots,
    IC_REDIST_DONORS2,     // used by alternate rounds)
    IC_COUNT
};
// FP64 counters (indices into DevBufs::dcnt)
enum {
    DC_CAND_MASS = 0,
    DC_CAND_VOLUME,
    DC_EXT_CAPACITY,
    DC_APPLIED_VOLUME,
    DC_REMOVAL_SCALE,
    DC_ADVECT_MASS,
    DC_COMPRESS_MASS,
    DC_REDUCE,  // scratch for lbm_reduce
    DC_COUNT = 12
};

constexpr int BED_BATCH = 8;              // moving-bed time levels computed per pair of launches
constexpr int BED_RING = BED_BATCH + 1;   // the batch never overwrites the level of the pending stream

struct DevBufs {
    int storage;
    size_t n, n2;
    void* f[2];  // populations, storage type, 19*n each; f[cur] is the state
    void* ftmp;  // arithmetic-type scratch of 19*n for materialised downloads
    void *rho, *ux, *uy, *uz, *omega, *nu_t, *smag;
    // bed time levels: a ring of `ring` slots, slot of step s = s % ring
    // (BED_RING for moving beds, whose levels are computed BED_BATCH steps
    // at a time; 2 for static beds)
    int ring;
    void *h[BED_RING], *uw[BED_RING], *vw[BED_RING], *ww[BED_RING], *eta9[BED_RING], *denom[BED_RING], *uw9[BED_RING];
    int* kband[BED_RING];
    int* ksolid[BED_RING];
    void* fsol[BED_RING];
    void* ibb[BED_RING];
    void* phi;  // phi of the latest bed time level (phi boundary-geometry modes only)
    void *piecewise, *u_inlet, *inflow;
    uint8_t* cellflags;  // bit0 lattice_open[0], bit1 near_obstacle (staged path, current slot)
    // water / VOF
    void *fill, *mass, *mass2, *mdelta, *excess, *excess2, *sa, *sb, *sc, *sd, *se, *smv, *fs_h;
    void *fs_qx, *fs_qy;  // column volume fluxes (height-function tracking)
    int *type, *type_prev, *excess_dir, *ext_gas, *obst, *obst_prev, *si;
    uint8_t* mdirty;  // VOF runs: cells whose populations changed after the stream's macro step
    int* rowdon;      // VOF runs: [2][nz * ny] rows holding a donor of a redistribution round
    int* rowact;      // VOF runs: [3][nz * ny] rows marked by the passes of a step (this step's and the last), rows with an isolated open cell
    uint8_t* tstate;  // free-surface runs: 2 type + open flag of every cell, for the stream (written by the macro step)
    int* icnt;      // [IC_COUNT]
    double* dcnt;   // [DC_COUNT]
    int* waveq;     // fused path, wavefront step: [0] error flag, [WAVE_FLAG0 + t] epoch of moments tile t
    double* rpart;  // FP64 partial sums of the reductions: 3 * reduce_chunks(n)
};
constexpr int WAVE_FLAG0 = 64;   // first tile flag (the error flag keeps a cache line of its own)
// tiles per plane of the fused kernels (tiles of `cells` cells, block_x in x)
inline size_t fused_tiles_per_plane(int nx, int ny, int block_x, int cells = 256) {
    const int bx = block_x < cells ? block_x : cells;
    const int by = cells / bx;
    return (size_t)((nx + bx - 1) / bx) * (size_t)((ny + by - 1) / by);
}
// ints of the wavefront flag buffer: room for the tiles of any block_x, of
// 256 cells (wavefront) and of 128 cells (tile-resident form)
inline size_t wave_flag_ints(int nx, int ny, int nz) {
    size_t m = 0;
    for (int cells = 128; cells <= 256; cells *= 2)
        for (int bx = 64; bx <= 256; bx *= 2) {
            const size_t t = fused_tiles_per_plane(nx, ny, bx, cells);
            m = t > m ? t : m;
        }
    return WAVE_FLAG0 + (size_t)nz * m;
}

// schedule of the fused step: wavefront (both passes in one launch, tiles
// in block-index order) or three launches; lag = planes of moments ahead of
// the collision, 0 for about as many tiles as the device holds at once
// (cus x 6 blocks: 12 planes at 256x128, 3 at 512x256, the measured best)
struct FusedSched {
    // FP32 storage, option "resident": the tile-resident form (measured
    // slower than the wavefront at 256x128x128: off by default)
    int resident = 0;
    int wavefront = 1;  // FP32 storage (the other storage types keep three launches)
    int lag = 0;
    int cus = 256;      // compute units of the device
    int strips = 1;     // XCD-aware tile order of the wavefront schedule (rows of tiles a multiple of 8)
    int epoch = 0;  // launches so far (the tile flags of launch e hold e)
};

// what the staged step may reuse from the step before (cleared by the C ABI
// whenever the geometry or any field is set from outside)
struct StagedState {
    int geo_ready = 0;    // cell flags and obstacle masks hold the (static) geometry
    int macro_ready = 0;  // rho and u hold the last stream's macro step (VOF runs, no conditioning)
    int band_ok = 0;      // the VOF row marks of the last step cover its end state (narrow band usable)
    int act_slot = 0;     // row-mark slot of the next step
    int band_allowed = 1;  // option vof_band (set by the C ABI before each step)
    int flood_sync = 1;    // option flood_sync (the same): the host stops the flood fill at its fixed point
};

// host-side launchers (lbm_kernels.hip)
// bed arrays of one time level at time t into `slot` (the previous level,
// for the bed velocity, is slot - 1 of the ring)
void launch_bed2d(const lbm_config_t& c, DevBufs& b, float t, int slot, hipStream_t s);
// moving bed: the levels of steps step0 .. step0 + nb - 1 (nb <= BED_BATCH)
// in two launches; the level of step0 - 1 must be current in its slot
void launch_bed2d_batch(const lbm_config_t& c, DevBufs& b, int64_t step0, int nb, hipStream_t s);
void launch_geometry_fields(const lbm_config_t& c, DevBufs& b, int slot, int32_t* lo, void* lf, int32_t* wt, int32_t* near,
                            void* phi, hipStream_t s);
void launch_cellflags(const lbm_config_t& c, DevBufs& b, int slot, hipStream_t s);
void launch_init_populations(const lbm_config_t& c, DevBufs& b, int slot, hipStream_t s);
// one staged step: bed arrays for slot must be current; advances b.f[cur]
// count_links: count the stream's link counters (the last step of a call)
void launch_staged_step(const lbm_config_t& c, DevBufs& b, int& cur, int slot, int64_t step, int omega_band_top, int count_links,
                        StagedState& st, hipStream_t s);
// fused step: mode 0 reads post-stream populations, mode 1 post-collision
// ones; slot_prev is the time level of the pending stream, slot the new one
// band_top: highest k at which any cell can have a cut or closed link
void launch_fused_step(const lbm_config_t& c, DevBufs& b, int& cur, int mode, int slot_prev, int slot, int64_t step,
                       int write_macro, int omega_band_top, int band_top, int block_x, FusedSched& fs, hipStream_t s);
// the wavefront step of several fused-path cases in one launch: the
// size of one case's table entry, its collision model (cases of one launch
// share the storage type and the model), the entry of one case's step
// (returns the case's blocks in each half of the grid, the moments half and
// the collision half; the entry is written to `out`, a host buffer of
// wave_case_bytes), and the launch over a device copy of the table (half:
// the blocks of one half)
size_t wave_case_bytes(int storage);
// the bed batches (launch_bed2d_batch) of several cases in one pair of
// launches: entry size, the entry of one case (its block counts in the two
// launches to na and nbb) and the launch over a device copy of the table
size_t bed_case_bytes(int storage);
void fill_bed_case(const lbm_config_t& c, DevBufs& b, int64_t step0, int nb, int a0, int b0, void* out, int* na, int* nbb);
void launch_bed_multi(int storage, const void* dev_table, int ncases, int na, int nbb, hipStream_t s);
// cases with equal keys can be advanced by one lbm_step_batch (-1: none)
int batch_key(lbm_handle_t* h);
// the argument table of a batch (device copy and pinned staging) and the
// event after its last copy (the staging buffer is free again)
struct BatchBuf {
    void* dev = nullptr;
    void* host = nullptr;
    size_t cap = 0;
    hipEvent_t copied = nullptr;
};
inline void batch_buf_free(BatchBuf& b) {
    if (b.copied) (void)hipEventSynchronize(b.copied);
    if (b.host) (void)hipHostFree(b.host);
    if (b.dev) (void)hipFree(b.dev);
    if (b.copied) (void)hipEventDestroy(b.copied);
    b = BatchBuf{};
}
// lbm_step_batch on stream bs with the table in buf; order = false: the
// caller keeps the handles' own streams free of work meanwhile
int step_batch(lbm_handle_t* const* hs, const int32_t* nsteps, int n, hipStream_t bs, BatchBuf& buf, bool order);
int wave_case_model(const lbm_config_t& c);
int fill_wave_case(const lbm_config_t& c, DevBufs& b, int cur, int mode, int slot_prev, int slot, int64_t step, int write_macro,
                   int omega_band_top, int band_top, int block_x, const FusedSched& fs, int block0, void* out);
void launch_wave_batch(int storage, int cm, const void* dev_table, int ncases, int half, hipStream_t s);
// an entry of fill_wave_case moved to another step: the fields that change
// from step to step (time levels, population buffers, mode, step, epoch,
// first block)
void patch_wave_case(int storage, void* entry, DevBufs& b, int cur, int mode, int slot_prev, int slot, int64_t step, int write_macro,
                     int epoch, int block0);
// post-collision state -> post-stream populations (arithmetic type) in dst
void launch_materialize(const lbm_config_t& c, DevBufs& b, int cur, int slot, void* dst, hipStream_t s);
// convert arithmetic-type populations in src into the storage buffer f[cur]
void launch_store_populations(const lbm_config_t& c, DevBufs& b, int cur, const void* src, int64_t step, hipStream_t s);
// deterministic FP64 reduction `what` of the state f[cur] (post_stream 0:
// a post-collision state, streamed on the fly with the time level `slot`)
// into out[0] (device), with `part` (at least reduce_chunks(n) doubles)
// for the partial sums
void launch_reduce(const lbm_config_t& c, DevBufs& b, int cur, int post_stream, int slot, int what, double* part, size_t max_parts,
                   double* out, hipStream_t s);
size_t reduce_chunks(size_t n);
// device address of the static partial-sum array of the current device
// (looked up once per device); *parts = its capacity in partials
double* noalloc_scratch(size_t* parts);
void launch_cut_link_count(const lbm_config_t& c, DevBufs& b, int slot, hipStream_t s);
bool fused_supported(const lbm_config_t& c);
// obstacle mask of a time level (init_fields, CODEX:2711-2712) into dst
void launch_obstacle(const lbm_config_t& c, DevBufs& b, int slot, int* dst, hipStream_t s);
// storage-type populations f[cur] -> arithmetic-type dst
void launch_load_populations(const lbm_config_t& c, DevBufs& b, int cur, void* dst, hipStream_t s);

}  // namespace lbm
