// lbm_capi.hip -- the extern "C" ABI of liblbm.so (declared in include/lbm.h).
//
// One handle = one case = one HIP stream.  The handle owns every device
// buffer; hosts pass plain pointers.  The population state is either the
// post-stream set f(n) of the reference (after initialisation, uploads and
// every step of the staged path) or the post-collision set f*(n) whose
// stream is still pending (after a step of the fused path); downloads and
// reductions of the populations apply the pending stream first, so callers
// always see the reference's f(n).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/lbm.h"
#include "lbm_internal.h"

using namespace lbm;

namespace {

std::string g_create_error;
std::atomic<int> g_live_handles{0};  // handles that exist in the process

struct Buf {
    void* p = nullptr;
};

}  // namespace

struct lbm_handle {
    lbm_config_t cfg{};
    DevBufs b{};
    hipStream_t stream = nullptr;
    // static (flat / piecewise) beds: the 2-D bed kernels run once per slot
    int slot_valid[BED_RING] = {};
    int cur = 0;
    int post_stream = 1;  // 1: f[cur] = f(step), 0: f[cur] = f*(step), stream pending
    int64_t step = 0;
    int path = LBM_PATH_STAGED;
    int omega_band_top = -1;
    int wall_top = 0;
    double piecewise_max = 0.0;
    int block_x = 64;  // x extent of a fused-kernel block (64, 128 or 256 cells; set from nx at creation)
    FusedSched sched;  // fused path: wavefront (one launch per step) or three launches
    StagedState staged;  // staged path: what a step may reuse from the previous one
    int vof_band = 1;    // option vof_band: VOF passes skip the rows settled far from the interface
    int flood_sync = 1;  // option flood_sync: the host reads the flood fill's flag every 16 sweeps
    int dev = 0;         // the device the handle was created on
    // option "resident" 2 (the default): the schedule of the fused step is
    // chosen by timing both on the handle's first long enough lbm_step call
    // (1 pending, 0 chosen or set explicitly)
    int sched_auto = 1;
    // fused path, option "graph": the launches of an lbm_step call are
    // captured into a HIP graph and launched as one; the executable graph is
    // kept and updated with the next call's launches (same kernels, new
    // arguments), and instantiated again when the launch sequence differs
    int graph = 0;
    hipGraphExec_t gexec = nullptr;
    // lbm_step_batch with this handle first: the table of every step's
    // arguments, and the ordering event
    BatchBuf batch;
    hipEvent_t ev_order = nullptr;  // this handle's stream has reached a batch boundary
    int initialized = 0;
    int geometry_done = 0;
    int timing = 0;
    double last_ms_step = 0.0, last_ms_kernel = 0.0;
    std::vector<hipEvent_t> ev;  // timing events: 2 per step + 2 around the loop
    std::string err;
    std::vector<void*> allocs;
    double* noalloc_part = nullptr;  // device address of the static reduction scratch (this handle's device)
    size_t noalloc_parts = 0;
};

#define CHECK(expr)                                                                                     \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess) throw std::runtime_error(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

static size_t real_size(const lbm_config_t& c) { return c.storage == LBM_STORE_FP64 ? 8 : 4; }
static size_t store_size(const lbm_config_t& c) {
    switch (c.storage) {
        case LBM_STORE_FP64: return 8;
        case LBM_STORE_FP16:
        case LBM_STORE_BF16: return 2;
        default: return 4;
    }
}

static const char* validate(const lbm_config_t* c) {
    if (!c) return "config is NULL";
    if (c->version != LBM_CONFIG_VERSION) return "config version mismatch";
    if (c->nx < 2 || c->ny < 2 || c->nz < 3) return "grid must be at least 2 x 2 x 3";
    if ((size_t)c->nx * c->ny * c->nz > (size_t)1 << 31) return "grid larger than 2^31 cells";
    if (c->storage < 0 || c->storage > 3) return "unknown storage";
    if (c->storage_shift < 0 || c->storage_shift > 24) return "storage_shift must be in [0, 24]";
    if (c->storage_rounding != LBM_ROUND_NEAREST && c->storage_rounding != LBM_ROUND_STOCHASTIC) return "unknown storage_rounding";
    for (int a = 0; a < 3; a++)
        if (!(fabsf(c->storage_u_ref[a]) <= 0.3f)) return "storage_u_ref components must lie in [-0.3, 0.3]";
    if ((c->physics_mode == LBM_PHYS_WATER || c->fs_enabled) && c->kernel_path == LBM_PATH_FUSED)
        return "the fused path covers air mode without a free surface only";
    if (!(c->tau0 > 0.5f)) return "tau0 must exceed 0.5";
    if (!(c->dx > 0.0f)) return "dx must be positive";
    if (c->vof_collapse_interval < 1) return "vof_collapse_interval must be positive";
    if (c->kernel_path == LBM_PATH_FUSED && !fused_supported(*c)) return "the fused path covers air mode without a free surface or the box obstacle";
    return nullptr;
}

static int resolve_path(const lbm_config_t& c) {
    if (c.kernel_path == LBM_PATH_STAGED) return LBM_PATH_STAGED;
    if (c.kernel_path == LBM_PATH_FUSED) return LBM_PATH_FUSED;
    return fused_supported(c) ? LBM_PATH_FUSED : LBM_PATH_STAGED;
}

// upper bound of the bed height over the whole run
static double bed_height_bound(const lbm_config_t& c, double piecewise_max) {
    double amp = 0.0;
    if (c.bed_profile == LBM_BED_MOVING_SINE) amp = 2.0 * fabs((double)c.bed_amp_m);
    if (c.bed_profile == LBM_BED_MOVING_SINE1) amp = fabs((double)c.bed_amp_m);
    if (c.bed_profile == LBM_BED_PIECEWISE) return piecewise_max;
    return (double)c.offset_m + amp;
}

// highest k at which a cell can switch between solid and fluid (moving beds)
static int band_top(const lbm_config_t& c) {
    if (!(c.bed_profile == LBM_BED_MOVING_SINE || c.bed_profile == LBM_BED_MOVING_SINE1) || !c.include_bed) return -1;
    int k = (int)ceil(bed_height_bound(c, 0.0) / (double)c.dx) + 2;
    return k >= c.nz ? c.nz - 1 : k;
}

// highest k at which any cell can have a cut or closed link: the per-column
// band of k_bed2d_b (ceil(h/dx + 1/2) + 1) plus one cell; every k for
// solid x or y walls
static int wall_band_top(const lbm_config_t& c, double piecewise_max) {
    if (c.x_boundary == LBM_XB_SOLID || c.y_boundary == LBM_YB_SOLID) return c.nz - 1;
    if (!c.include_bed) return -1;
    double h = bed_height_bound(c, piecewise_max);
    double k = ceil(h / (double)c.dx + 0.5) + 2.0;
    if (k < -1.0) return -1;
    return k >= (double)(c.nz - 1) ? c.nz - 1 : (int)k;
}

// slots of the bed ring: a moving bed keeps BED_RING time levels (computed
// BED_BATCH at a time), a static one (or the phi modes, one phi field per
// level) two
static int bed_ring(const lbm_config_t& c) {
    const bool moving = c.bed_profile == LBM_BED_MOVING_SINE || c.bed_profile == LBM_BED_MOVING_SINE1;
    return moving && !c.phi_fraction ? BED_RING : 2;
}

static size_t device_bytes(const lbm_config_t& c, int path) {
    size_t n = (size_t)c.nx * c.ny * c.nz, n2 = (size_t)c.nx * c.ny;
    size_t R = real_size(c);
    size_t bytes = 2 * 19 * n * store_size(c);  // ping-pong populations
    bytes += 5 * n * R + n;                      // rho, u, omega, cell flags
    if (path == LBM_PATH_STAGED) bytes += 2 * n * R + 2 * n * 4;  // nu_t, smag, obstacle masks
    bytes += (size_t)bed_ring(c) * n2 * (78 * R + 8) + (c.nx + c.nz + (size_t)c.nz * c.ny) * R;
    if (c.fs_enabled) bytes += n * (12 * R + 7 * 4 + 2) + 6 * n * R + 3 * n2 * R + 20 * (size_t)c.nz * c.ny;
    if (c.phi_fraction) bytes += n * R;
    return bytes + (1 << 20);
}

static void* dalloc(lbm_handle* h, size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) throw std::runtime_error(std::string("hipMalloc(") + std::to_string(bytes) + "): " + hipGetErrorString(e));
    h->allocs.push_back(p);
    return p;
}

static void free_all(lbm_handle* h) {
    for (void* p : h->allocs) (void)hipFree(p);
    h->allocs.clear();
}

static void fill_real(lbm_handle* h, void* dst, size_t count, double value) {
    size_t R = real_size(h->cfg);
    std::vector<char> tmp(count * R);
    if (R == 8)
        for (size_t i = 0; i < count; i++) ((double*)tmp.data())[i] = value;
    else
        for (size_t i = 0; i < count; i++) ((float*)tmp.data())[i] = (float)value;
    CHECK(hipMemcpyAsync(dst, tmp.data(), tmp.size(), hipMemcpyHostToDevice, h->stream));
    CHECK(hipStreamSynchronize(h->stream));
}

static void allocate(lbm_handle* h) {
    const lbm_config_t& c = h->cfg;
    DevBufs& b = h->b;
    size_t n = (size_t)c.nx * c.ny * c.nz, n2 = (size_t)c.nx * c.ny;
    size_t R = real_size(c), Sz = store_size(c);
    b.storage = c.storage;
    b.n = n;
    b.n2 = n2;
    b.f[0] = dalloc(h, 19 * n * Sz);
    b.f[1] = dalloc(h, 19 * n * Sz);
    CHECK(hipMemsetAsync(b.f[0], 0, 19 * n * Sz, h->stream));
    CHECK(hipMemsetAsync(b.f[1], 0, 19 * n * Sz, h->stream));
    b.ftmp = nullptr;  // allocated on first download
    b.rho = dalloc(h, n * R);
    b.ux = dalloc(h, n * R);
    b.uy = dalloc(h, n * R);
    b.uz = dalloc(h, n * R);
    b.omega = dalloc(h, n * R);
    b.cellflags = (uint8_t*)dalloc(h, n);
    if (h->path == LBM_PATH_STAGED) {
        b.nu_t = dalloc(h, n * R);
        b.smag = dalloc(h, n * R);
        b.obst = (int*)dalloc(h, n * 4);
        b.obst_prev = (int*)dalloc(h, n * 4);
        CHECK(hipMemsetAsync(b.nu_t, 0, n * R, h->stream));
        CHECK(hipMemsetAsync(b.smag, 0, n * R, h->stream));
        CHECK(hipMemsetAsync(b.obst, 0, n * 4, h->stream));
        CHECK(hipMemsetAsync(b.obst_prev, 0, n * 4, h->stream));
    } else {
        b.obst = (int*)dalloc(h, n * 4);  // init writes the t=0 mask here
        b.obst_prev = nullptr;
        // error flag + one flag per moments tile
        const size_t wq = wave_flag_ints(c.nx, c.ny, c.nz);
        b.waveq = (int*)dalloc(h, wq * sizeof(int));
        CHECK(hipMemsetAsync(b.waveq, 0, wq * sizeof(int), h->stream));
    }
    for (int s = 0; s < b.ring; s++) {
        b.h[s] = dalloc(h, n2 * R);
        b.uw[s] = dalloc(h, n2 * R);
        b.vw[s] = dalloc(h, n2 * R);
        b.ww[s] = dalloc(h, n2 * R);
        b.eta9[s] = dalloc(h, 9 * n2 * R);
        b.uw9[s] = dalloc(h, 27 * n2 * R);
        b.denom[s] = dalloc(h, n2 * R);
        b.kband[s] = (int*)dalloc(h, n2 * 4);
        b.ksolid[s] = (int*)dalloc(h, n2 * 4);
        b.fsol[s] = dalloc(h, 19 * n2 * R);
        b.ibb[s] = dalloc(h, 18 * n2 * R);
        CHECK(hipMemsetAsync(b.uw[s], 0, n2 * R, h->stream));
        CHECK(hipMemsetAsync(b.vw[s], 0, n2 * R, h->stream));
        CHECK(hipMemsetAsync(b.ww[s], 0, n2 * R, h->stream));
    }
    if (c.phi_fraction) b.phi = dalloc(h, n * R);
    b.piecewise = dalloc(h, c.nx * R);
    b.u_inlet = dalloc(h, c.nz * R);
    b.inflow = dalloc(h, (size_t)c.nz * c.ny * R);
    CHECK(hipMemsetAsync(b.u_inlet, 0, c.nz * R, h->stream));
    CHECK(hipMemsetAsync(b.inflow, 0, (size_t)c.nz * c.ny * R, h->stream));
    b.icnt = (int*)dalloc(h, IC_COUNT * sizeof(int));
    b.dcnt = (double*)dalloc(h, DC_COUNT * sizeof(double));
    CHECK(hipMemsetAsync(b.icnt, 0, IC_COUNT * sizeof(int), h->stream));
    CHECK(hipMemsetAsync(b.dcnt, 0, DC_COUNT * sizeof(double), h->stream));
    b.rpart = (double*)dalloc(h, 3 * reduce_chunks(n) * sizeof(double));
    if (c.fs_enabled) {
        b.fill = dalloc(h, n * R);
        b.mass = dalloc(h, n * R);
        b.mass2 = dalloc(h, n * R);
        b.mdelta = dalloc(h, n * R);
        b.excess = dalloc(h, n * R);
        b.excess2 = dalloc(h, n * R);
        b.sa = dalloc(h, n * R);
        b.sb = dalloc(h, n * R);
        b.sc = dalloc(h, n * R);
        b.sd = dalloc(h, n * R);
        b.se = dalloc(h, n * R);
        b.smv = dalloc(h, 6 * n * R);
        b.fs_h = dalloc(h, n2 * R);
        b.fs_qx = dalloc(h, n2 * R);
        b.fs_qy = dalloc(h, n2 * R);
        b.type = (int*)dalloc(h, n * 4);
        b.type_prev = (int*)dalloc(h, n * 4);
        b.excess_dir = (int*)dalloc(h, n * 4);
        b.ext_gas = (int*)dalloc(h, n * 4);
        b.si = (int*)dalloc(h, n * 4);
        b.tstate = (uint8_t*)dalloc(h, n);
        b.mdirty = (uint8_t*)dalloc(h, n);
        CHECK(hipMemsetAsync(b.mdirty, 0, n, h->stream));
        b.rowdon = (int*)dalloc(h, 2 * (size_t)c.nz * c.ny * 4);
        CHECK(hipMemsetAsync(b.rowdon, 0, 2 * (size_t)c.nz * c.ny * 4, h->stream));
        b.rowact = (int*)dalloc(h, 3 * (size_t)c.nz * c.ny * 4);
        CHECK(hipMemsetAsync(b.rowact, 0, 3 * (size_t)c.nz * c.ny * 4, h->stream));
        CHECK(hipMemsetAsync(b.mdelta, 0, n * R, h->stream));
        CHECK(hipMemsetAsync(b.excess, 0, n * R, h->stream));
        CHECK(hipMemsetAsync(b.excess_dir, 0, n * 4, h->stream));
        CHECK(hipMemsetAsync(b.ext_gas, 0, n * 4, h->stream));
        fill_real(h, b.fill, n, 1.0);
        fill_real(h, b.mass, n, (double)c.rho0);
        fill_real(h, b.fs_h, n2, (double)c.free_surface_base_m);
        std::vector<int> liquid(n, LBM_FS_LIQUID);
        CHECK(hipMemcpyAsync(b.type, liquid.data(), n * 4, hipMemcpyHostToDevice, h->stream));
        CHECK(hipMemcpyAsync(b.type_prev, liquid.data(), n * 4, hipMemcpyHostToDevice, h->stream));
    }
    // init_constants (CODEX:1022-1063): rho = rho0, u = 0, omega = 1/tau0,
    // bed height = the bed level
    fill_real(h, b.rho, n, (double)c.rho0);
    CHECK(hipMemsetAsync(b.ux, 0, n * R, h->stream));
    CHECK(hipMemsetAsync(b.uy, 0, n * R, h->stream));
    CHECK(hipMemsetAsync(b.uz, 0, n * R, h->stream));
    double om0 = R == 8 ? 1.0 / (double)c.tau0 : (double)(1.0f / c.tau0);
    fill_real(h, b.omega, n, om0);
    for (int sl = 0; sl < b.ring; sl++) fill_real(h, b.h[sl], n2, (double)c.offset_m);
    fill_real(h, b.piecewise, c.nx, (double)c.offset_m);
    CHECK(hipStreamSynchronize(h->stream));
}

static void ensure_ftmp(lbm_handle* h) {
    if (!h->b.ftmp) h->b.ftmp = dalloc(h, 19 * h->b.n * real_size(h->cfg));
}

// bed time level (slot of the ring) of a step
static int slot_of(const lbm_handle* h, int64_t step) { return (int)(step % h->b.ring); }

// the reference's post-stream populations f(step) as an arithmetic-type
// device buffer (either the state itself or ftmp)
static const void* post_stream_populations(lbm_handle* h) {
    if (h->post_stream) {
        if (h->cfg.storage == LBM_STORE_FP32 || h->cfg.storage == LBM_STORE_FP64) return h->b.f[h->cur];
        ensure_ftmp(h);
        launch_load_populations(h->cfg, h->b, h->cur, h->b.ftmp, h->stream);
        return h->b.ftmp;
    }
    ensure_ftmp(h);
    launch_materialize(h->cfg, h->b, h->cur, slot_of(h, h->step), h->b.ftmp, h->stream);
    return h->b.ftmp;
}

#define API_BEGIN(h)                      \
    if (!(h)) return LBM_E_ARG;           \
    try {
#define API_END(h)                                   \
    }                                                \
    catch (const std::exception& e) {                \
        (h)->err = e.what();                         \
        return LBM_E_HIP;                            \
    }                                                \
    return LBM_OK;

static int fail(lbm_handle* h, int code, const std::string& msg) {
    h->err = msg;
    return code;
}

// the wavefront step's bounded waits leave a flag when one ran out (the
// step's results are then not valid); read and cleared after the stream
// has drained
static void check_wave_flag(lbm_handle* h) {
    if (!h->b.waveq) return;
    int flag = 0;
    CHECK(hipMemcpyAsync(&flag, h->b.waveq, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    CHECK(hipStreamSynchronize(h->stream));
    if (flag) {
        int zero = 0;
        CHECK(hipMemcpyAsync(h->b.waveq, &zero, sizeof(int), hipMemcpyHostToDevice, h->stream));
        CHECK(hipStreamSynchronize(h->stream));
        throw std::runtime_error("wavefront step: a wait for the moments of a neighbouring plane ran out; the state is not valid");
    }
}

namespace lbm {
// cases with the same key can share a batch (fused path, one storage type
// and collision model); -1 for a case that cannot be batched
int batch_key(lbm_handle_t* h) {
    if (!h || h->path != LBM_PATH_FUSED || !h->initialized) return -1;
    return h->cfg.storage * 4 + wave_case_model(h->cfg);
}
}  // namespace lbm

// Launches whose blocks wait for other blocks of the same launch (the
// fused step's wavefront and tile-resident schedules, the merged grid of
// lbm_step_batch): a block waits only for blocks the dispatcher started
// before it, which is safe while no other such launch runs on the device;
// two running at once can each fill the compute units that the other's
// unstarted blocks need.  While more than one handle exists in the process,
// these launches are chained on one event per device: each starts after the
// previous one ended (other kernels still run beside it).  Handles with the
// option graph and separate processes stepping on one GPU are not chained:
// they should use the three-launch schedule (options wavefront 0, resident
// 0) when they step at the same time as others.
static std::mutex g_wait_m;
static hipEvent_t g_wait_ev[16];
template <class F>
static void chain_waiting(int dev, hipStream_t s, bool waits, F&& launch) {
    if (!waits || g_live_handles.load() < 2) {
        launch();
        return;
    }
    std::lock_guard<std::mutex> g(g_wait_m);
    hipEvent_t& ev = g_wait_ev[dev & 15];
    if (!ev) CHECK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    CHECK(hipStreamWaitEvent(s, ev, 0));
    launch();
    CHECK(hipEventRecord(ev, s));
}

extern "C" {

int lbm_create(const lbm_config_t* cfg, lbm_handle_t** out) {
    if (!out) {
        g_create_error = "out is NULL";
        return LBM_E_ARG;
    }
    *out = nullptr;
    const char* why = validate(cfg);
    if (why) {
        g_create_error = why;
        return LBM_E_CONFIG;
    }
    lbm_handle* h = new lbm_handle();
    h->cfg = *cfg;
    h->path = resolve_path(*cfg);
    h->block_x = cfg->nx >= 256 ? 256 : (cfg->nx >= 128 ? 128 : 64);
    h->omega_band_top = band_top(*cfg);
    h->piecewise_max = (double)cfg->offset_m;
    h->wall_top = wall_band_top(*cfg, h->piecewise_max);
    h->b.ring = bed_ring(*cfg);
    try {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw std::runtime_error("no HIP device visible");
        CHECK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        int dev = 0, cus = 0;
        CHECK(hipGetDevice(&dev));
        h->dev = dev;
        CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        h->sched.cus = cus > 0 ? cus : 256;
        allocate(h);
        // the static scratch of lbm_reduce_noalloc: its address on this
        // device, looked up here so that the entry point itself does nothing
        // but launch
        h->noalloc_part = noalloc_scratch(&h->noalloc_parts);
    } catch (const std::exception& e) {
        g_create_error = e.what();
        free_all(h);
        if (h->stream) (void)hipStreamDestroy(h->stream);
        delete h;
        return LBM_E_NOMEM;
    }
    *out = h;
    g_live_handles.fetch_add(1);
    return LBM_OK;
}

void lbm_destroy(lbm_handle_t* h) {
    if (!h) return;
    g_live_handles.fetch_sub(1);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    batch_buf_free(h->batch);
    if (h->ev_order) (void)hipEventDestroy(h->ev_order);
    for (hipEvent_t e : h->ev) (void)hipEventDestroy(e);
    free_all(h);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* lbm_last_error(lbm_handle_t* h) { return h ? h->err.c_str() : g_create_error.c_str(); }

int lbm_geometry(lbm_handle_t* h, float t) {
    API_BEGIN(h)
    launch_bed2d(h->cfg, h->b, t, 0, h->stream);
    for (int sl = 0; sl < BED_RING; sl++) h->slot_valid[sl] = sl == 0;
    h->staged = StagedState{};
    launch_obstacle(h->cfg, h->b, 0, h->b.obst, h->stream);
    h->geometry_done = 1;
    h->step = 0;
    CHECK(hipStreamSynchronize(h->stream));
    API_END(h)
}

int lbm_init_populations(lbm_handle_t* h) {
    API_BEGIN(h)
    if (!h->geometry_done) return fail(h, LBM_E_STATE, "lbm_geometry must run before lbm_init_populations");
    h->cur = 0;
    h->staged = StagedState{};
    launch_init_populations(h->cfg, h->b, 0, h->stream);
    h->post_stream = 1;
    h->step = 0;
    h->initialized = 1;
    CHECK(hipStreamSynchronize(h->stream));
    API_END(h)
}

int lbm_initialize(lbm_handle_t* h, const lbm_init_params_t* p) {
    if (!h) return LBM_E_ARG;
    if (!p) return fail(h, LBM_E_ARG, "params is NULL");
    if (p->version != LBM_INIT_PARAMS_VERSION) return fail(h, LBM_E_CONFIG, "init params version mismatch");
    const lbm_config_t& c = h->cfg;
    const size_t n = h->b.n, n2 = h->b.n2, R = real_size(c);
    int rc = lbm_geometry(h, 0.0f);
    if (rc) return rc;
    std::vector<int32_t> obst(n), type(c.fs_enabled ? n : 0);
    std::vector<char> bed(n2 * R), rho(n * R), ux(n * R), uy(n * R), uz(n * R), uin(c.nz * R), infl((size_t)c.nz * c.ny * R);
    std::vector<char> fill(c.fs_enabled ? n * R : 0), fsh(c.fs_enabled ? n2 * R : 0);
    if ((rc = lbm_download_field(h, LBM_FIELD_OBSTACLE, obst.data(), n * 4))) return rc;
    if ((rc = lbm_download_field(h, LBM_FIELD_WATER_H, bed.data(), n2 * R))) return rc;
    rc = lbm_initial_fields(&c, p, obst.data(), bed.data(), rho.data(), ux.data(), uy.data(), uz.data(), uin.data(), infl.data(),
                            c.fs_enabled ? fill.data() : nullptr, c.fs_enabled ? type.data() : nullptr,
                            c.fs_enabled ? fsh.data() : nullptr);
    if (rc) return fail(h, rc, "lbm_initial_fields failed");
    struct Up {
        int32_t f;
        const void* d;
        size_t b;
    } ups[] = {{LBM_FIELD_U_INLET_PROFILE, uin.data(), uin.size()}, {LBM_FIELD_INFLOW_UX, infl.data(), infl.size()},
               {LBM_FIELD_RHO, rho.data(), rho.size()},          {LBM_FIELD_UX, ux.data(), ux.size()},
               {LBM_FIELD_UY, uy.data(), uy.size()},             {LBM_FIELD_UZ, uz.data(), uz.size()}};
    for (const Up& u : ups)
        if ((rc = lbm_upload_field(h, u.f, u.d, u.b))) return rc;
    if (c.fs_enabled) {
        if ((rc = lbm_upload_field(h, LBM_FIELD_FS_FILL, fill.data(), fill.size()))) return rc;
        if ((rc = lbm_upload_field(h, LBM_FIELD_FS_TYPE, type.data(), n * 4))) return rc;
        if ((rc = lbm_upload_field(h, LBM_FIELD_FS_TYPE_PREV, type.data(), n * 4))) return rc;
        if ((rc = lbm_upload_field(h, LBM_FIELD_FS_H, fsh.data(), fsh.size()))) return rc;
    }
    return lbm_init_populations(h);
}

// enqueue the launches of nsteps steps on the handle's stream
static void enqueue_steps(lbm_handle* h, int32_t nsteps) {
    const lbm_config_t& c = h->cfg;
    if (h->timing) CHECK(hipEventRecord(h->ev[0], h->stream));
    const bool fixed = c.bed_profile == LBM_BED_FLAT || c.bed_profile == LBM_BED_PIECEWISE;
    const bool batched = h->b.ring == BED_RING;  // a moving bed: its levels come BED_BATCH steps at a time
    const int64_t last_step = h->step + nsteps;
    int64_t bed_ready = h->step;  // the level of the current step is current
    for (int s = 0; s < nsteps; s++) {
        int64_t step = h->step + 1;
        float t = (float)((double)(step + 1) * c.dt_phys);  // CODEX:3874
        int slot = slot_of(h, step);
        if (batched) {
            if (step > bed_ready) {
                const int nb = (int)std::min<int64_t>(BED_BATCH, last_step - step + 1);
                launch_bed2d_batch(c, h->b, step, nb, h->stream);
                bed_ready = step + nb - 1;
            }
        } else if (!(fixed && h->slot_valid[slot])) {  // a static bed is the same at every time level
            launch_bed2d(c, h->b, t, slot, h->stream);
            h->slot_valid[slot] = 1;
        }
        if (h->timing) CHECK(hipEventRecord(h->ev[2 + 2 * s], h->stream));
        if (h->path == LBM_PATH_STAGED) {
            if (!h->post_stream) throw std::runtime_error("internal: staged step on a post-collision state");
            h->staged.band_allowed = h->vof_band;
            h->staged.flood_sync = h->flood_sync;
            launch_staged_step(c, h->b, h->cur, slot, step, h->omega_band_top, s == nsteps - 1, h->staged, h->stream);
        } else {
            int last = s == nsteps - 1;
            if (h->sched.epoch >= INT32_MAX - 1) {
                // the tile flags hold launch numbers: start them over
                const size_t wq = wave_flag_ints(c.nx, c.ny, c.nz);
                CHECK(hipMemsetAsync(h->b.waveq + WAVE_FLAG0, 0, (wq - WAVE_FLAG0) * sizeof(int), h->stream));
                h->sched.epoch = 0;
            }
            h->sched.epoch++;
            // (a handle with the option graph is not chained: its stream is
            // captured, and an event recorded on a stream that begins a
            // capture cannot be waited on from another)
            // (16-bit storage waits only on the tile-resident grid, FP64 never)
            const bool waits = c.storage == LBM_STORE_FP32 ? (h->sched.wavefront || h->sched.resident)
                                                           : (c.storage != LBM_STORE_FP64 && h->sched.resident);
            chain_waiting(h->dev, h->stream, waits && !h->graph, [&] {
                launch_fused_step(c, h->b, h->cur, h->post_stream ? 0 : 1, slot_of(h, step - 1), slot, step, last, h->omega_band_top,
                                  h->wall_top, h->block_x, h->sched, h->stream);
            });
            h->post_stream = 0;
        }
        if (h->timing) CHECK(hipEventRecord(h->ev[3 + 2 * s], h->stream));
        h->step = step;
    }
}

// the same launches captured into a graph (fused path only: its steps are
// kernel launches alone) and launched as one
static void graph_steps(lbm_handle* h, int32_t nsteps) {
    if (h->sched.epoch >= INT32_MAX - 1 - nsteps) {  // the flag reset (a fill) stays outside the graph
        const size_t wq = wave_flag_ints(h->cfg.nx, h->cfg.ny, h->cfg.nz);
        CHECK(hipMemsetAsync(h->b.waveq + WAVE_FLAG0, 0, (wq - WAVE_FLAG0) * sizeof(int), h->stream));
        h->sched.epoch = 0;
    }
    // the handle's state advances while the launches are recorded: keep it
    // to restore if the capture fails
    const int64_t step0 = h->step;
    const int cur0 = h->cur, post0 = h->post_stream, epoch0 = h->sched.epoch;
    int valid0[BED_RING];
    memcpy(valid0, h->slot_valid, sizeof valid0);
    CHECK(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipGraph_t g = nullptr;
    try {
        enqueue_steps(h, nsteps);
    } catch (...) {
        (void)hipStreamEndCapture(h->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        h->step = step0;
        h->cur = cur0;
        h->post_stream = post0;
        h->sched.epoch = epoch0;
        memcpy(h->slot_valid, valid0, sizeof valid0);
        throw;
    }
    CHECK(hipStreamEndCapture(h->stream, &g));
    bool updated = false;
    if (h->gexec) {
        hipGraphNode_t err_node = nullptr;
        hipGraphExecUpdateResult res;
        updated = hipGraphExecUpdate(h->gexec, g, &err_node, &res) == hipSuccess && res == hipGraphExecUpdateSuccess;
        if (!updated) {
            (void)hipGetLastError();
            (void)hipGraphExecDestroy(h->gexec);
            h->gexec = nullptr;
        }
    }
    hipError_t e = hipSuccess;
    if (!updated) e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipGraphLaunch(h->gexec, h->stream);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) throw std::runtime_error(std::string("graph launch: ") + hipGetErrorString(e));
}

// The fused step's two schedules, the wavefront (both passes in one
// launch, tiles in block order) and the tile-resident grid, give the same
// state bit for bit; which is faster depends on the grid and the collision
// model (neither wins everywhere).  On the first lbm_step call of at least
// AUTO_MIN steps, after one untimed step of each, AUTO_RUN steps of each are
// timed twice with HIP events on the handle's stream, and the faster is kept
// (the call waits for the device meanwhile).
constexpr int AUTO_RUN = 4;
constexpr int AUTO_MIN = 2 + 4 * AUTO_RUN;
static void auto_schedule(lbm_handle* h, int32_t& nsteps) {
    h->sched_auto = 0;
    hipEvent_t e[2] = {nullptr, nullptr};
    float best[2] = {1e30f, 1e30f};
    try {
        CHECK(hipEventCreate(&e[0]));
        CHECK(hipEventCreate(&e[1]));
        auto run = [&](int res, int ns) {
            h->sched.resident = res;
            CHECK(hipEventRecord(e[0], h->stream));
            enqueue_steps(h, ns);
            CHECK(hipEventRecord(e[1], h->stream));
            CHECK(hipEventSynchronize(e[1]));
            float ms = 0.0f;
            CHECK(hipEventElapsedTime(&ms, e[0], e[1]));
            return ms;
        };
        run(0, 1);
        run(1, 1);
        for (int r = 0; r < 2; r++)
            for (int res = 0; res < 2; res++) best[res] = std::min(best[res], run(res, AUTO_RUN));
    } catch (...) {
        h->sched.resident = 0;
        for (hipEvent_t x : e)
            if (x) (void)hipEventDestroy(x);
        throw;
    }
    for (hipEvent_t x : e) (void)hipEventDestroy(x);
    h->sched.resident = best[1] < best[0] ? 1 : 0;
    nsteps -= AUTO_MIN;
}

int lbm_step(lbm_handle_t* h, int32_t nsteps) {
    API_BEGIN(h)
    if (!h->initialized) return fail(h, LBM_E_STATE, "lbm_init_populations (or an upload of F) must run first");
    if (nsteps < 0) return fail(h, LBM_E_ARG, "nsteps must be >= 0");
    if (nsteps == 0) return LBM_OK;
    if (h->sched_auto && h->path == LBM_PATH_FUSED && h->cfg.storage != LBM_STORE_FP64 && !h->timing && nsteps >= AUTO_MIN) {
        auto_schedule(h, nsteps);
        if (nsteps == 0) return LBM_OK;
    }
    size_t need = h->timing ? (size_t)2 * nsteps + 2 : 0;
    while (h->ev.size() < need) {
        hipEvent_t e;
        CHECK(hipEventCreate(&e));
        h->ev.push_back(e);
    }
    if (h->graph && h->path == LBM_PATH_FUSED && !h->timing && nsteps > 1)
        graph_steps(h, nsteps);
    else
        enqueue_steps(h, nsteps);
    if (h->timing) {
        CHECK(hipEventRecord(h->ev[1], h->stream));
        CHECK(hipEventSynchronize(h->ev[1]));
        float ms = 0.0f;
        CHECK(hipEventElapsedTime(&ms, h->ev[0], h->ev[1]));
        h->last_ms_step = ms / nsteps;
        double k = 0.0;
        for (int s = 0; s < nsteps; s++) {
            float m = 0.0f;
            CHECK(hipEventElapsedTime(&m, h->ev[2 + 2 * s], h->ev[3 + 2 * s]));
            k += m;
        }
        h->last_ms_kernel = k / nsteps;
    }
    API_END(h)
}

}  // extern "C"

namespace lbm {
// several fused-path handles advanced together on stream bs: one launch
// per step runs the wavefront step of every handle that still has steps
// (its tiles in one grid, each case's arguments from a table in buf); the
// bed levels of each case are launched as in lbm_step.  order: wait for
// every handle's earlier work on its own stream first, and make its later
// work wait for the batch (otherwise the caller keeps the handles' streams
// free of work while they are batched).  Errors are reported on hs[0].
int step_batch(lbm_handle* const* hs, const int32_t* nsteps, int n, hipStream_t bs, BatchBuf& buf, bool order) {
    lbm_handle* h0 = hs[0];
    const int storage = hs[0]->cfg.storage;
    const int cm = wave_case_model(hs[0]->cfg);
    int32_t smax = 0;
    for (int i = 0; i < n; i++) {
        lbm_handle* h = hs[i];
        if (!h) return fail(h0, LBM_E_ARG, "a handle is NULL");
        if (!h->initialized) return fail(h0, LBM_E_STATE, "every handle must be initialised");
        if (h->path != LBM_PATH_FUSED) return fail(h0, LBM_E_ARG, "lbm_step_batch runs fused-path cases only");
        if (h->cfg.storage != storage || wave_case_model(h->cfg) != cm)
            return fail(h0, LBM_E_ARG, "the cases of one batch share the storage type and the collision model");
        if (nsteps[i] < 0) return fail(h0, LBM_E_ARG, "nsteps must be >= 0");
        for (int k = 0; k < i; k++)
            if (hs[k] == h) return fail(h0, LBM_E_ARG, "a handle appears twice in the batch");
        smax = std::max(smax, nsteps[i]);
    }
    if (smax == 0) return LBM_OK;
    const size_t E = wave_case_bytes(storage), EB = bed_case_bytes(storage);
    size_t entries = 0, bed_entries = 0;
    for (int i = 0; i < n; i++) {
        entries += (size_t)nsteps[i];
        if (hs[i]->b.ring == BED_RING) bed_entries += ((size_t)nsteps[i] + BED_BATCH - 1) / BED_BATCH;
    }
    const size_t bed_base = (entries * E + 255) / 256 * 256;  // the bed tables after the step tables
    const size_t need = bed_base + bed_entries * EB;
    if (!buf.copied) CHECK(hipEventCreateWithFlags(&buf.copied, hipEventDisableTiming));
    CHECK(hipEventSynchronize(buf.copied));  // the staging buffer is free again
    if (need > buf.cap) {
        CHECK(hipStreamSynchronize(bs));  // the old table may still be read
        if (buf.host) CHECK(hipHostFree(buf.host));
        if (buf.dev) CHECK(hipFree(buf.dev));
        buf.host = buf.dev = nullptr;
        buf.cap = 0;
        CHECK(hipHostMalloc(&buf.host, need, hipHostMallocDefault));
        CHECK(hipMalloc(&buf.dev, need));
        buf.cap = need;
    }
    // every handle's earlier work comes first
    for (int i = 0; i < n; i++) {
        lbm_handle* h = hs[i];
        if (order && h->stream != bs) {
            if (!h->ev_order) CHECK(hipEventCreateWithFlags(&h->ev_order, hipEventDisableTiming));
            CHECK(hipEventRecord(h->ev_order, h->stream));
            CHECK(hipStreamWaitEvent(bs, h->ev_order, 0));
        }
        if (h->sched.epoch >= INT32_MAX - 1 - nsteps[i]) {  // the tile flags hold launch numbers: start them over
            const size_t wq = wave_flag_ints(h->cfg.nx, h->cfg.ny, h->cfg.nz);
            CHECK(hipMemsetAsync(h->b.waveq + WAVE_FLAG0, 0, (wq - WAVE_FLAG0) * sizeof(int), bs));
            h->sched.epoch = 0;
        }
    }
    // the table: for step s, the entries of the cases that still step; each
    // case's entry is built once (its constants) and moved to every step
    std::vector<char> tmpl((size_t)n * E);
    std::vector<int> half_of(n);
    for (int i = 0; i < n; i++) {
        lbm_handle* h = hs[i];
        half_of[i] = fill_wave_case(h->cfg, h->b, h->cur, 1, slot_of(h, h->step), slot_of(h, h->step + 1), h->step + 1, 0,
                                    h->omega_band_top, h->wall_top, h->block_x, h->sched, 0, tmpl.data() + (size_t)i * E);
    }
    std::vector<size_t> off(smax), bed_off(smax);
    std::vector<int> cnt(smax), blocks(smax), bed_cnt(smax), bed_na(smax), bed_nb(smax);
    {
        // moving beds: the levels of BED_BATCH steps at a time, as lbm_step
        // computes them, for every case that needs them at step s in one
        // pair of launches before the step
        std::vector<int64_t> ready(n);
        for (int i = 0; i < n; i++) ready[i] = hs[i]->step;
        size_t eb = 0;
        for (int32_t s = 0; s < smax; s++) {
            bed_off[s] = eb;
            int na = 0, nbb = 0, c = 0;
            for (int i = 0; i < n; i++) {
                lbm_handle* h = hs[i];
                if (s >= nsteps[i] || h->b.ring != BED_RING) continue;
                const int64_t step = h->step + 1 + s;
                if (step <= ready[i]) continue;
                const int nb = (int)std::min<int64_t>(BED_BATCH, (int64_t)nsteps[i] - s);
                int ca = 0, cb = 0;
                fill_bed_case(h->cfg, h->b, step, nb, na, nbb, (char*)buf.host + bed_base + eb * EB, &ca, &cb);
                na += ca;
                nbb += cb;
                ready[i] = step + nb - 1;
                eb++;
                c++;
            }
            bed_cnt[s] = c;
            bed_na[s] = na;
            bed_nb[s] = nbb;
        }
        size_t e = 0;
        for (int32_t s = 0; s < smax; s++) {
            off[s] = e;
            int block0 = 0, c = 0;
            for (int i = 0; i < n; i++) {
                lbm_handle* h = hs[i];
                if (s >= nsteps[i]) continue;
                const int64_t step = h->step + 1 + s;
                const int cur = (h->cur + s) & 1;
                const int mode = (s == 0 && h->post_stream) ? 0 : 1;
                char* ent = (char*)buf.host + e * E;
                memcpy(ent, tmpl.data() + (size_t)i * E, E);
                patch_wave_case(storage, ent, h->b, cur, mode, slot_of(h, step - 1), slot_of(h, step), step, s == nsteps[i] - 1,
                                h->sched.epoch + 1 + s, block0);
                block0 += half_of[i];
                e++;
                c++;
            }
            cnt[s] = c;
            blocks[s] = block0;
        }
    }
    CHECK(hipMemcpyAsync(buf.dev, buf.host, need, hipMemcpyHostToDevice, bs));
    CHECK(hipEventRecord(buf.copied, bs));
    // the launches: the bed levels (moving beds: the batches of this step
    // in one pair of launches; static beds: once per slot, as lbm_step), then
    // the step of every case
    for (int32_t s = 0; s < smax; s++) {
        if (bed_cnt[s]) launch_bed_multi(storage, (const char*)buf.dev + bed_base + bed_off[s] * EB, bed_cnt[s], bed_na[s], bed_nb[s], bs);
        for (int i = 0; i < n; i++) {
            lbm_handle* h = hs[i];
            if (s >= nsteps[i] || h->b.ring == BED_RING) continue;
            const lbm_config_t& c = h->cfg;
            const int64_t step = h->step + 1;
            const int slot = slot_of(h, step);
            const bool fixed = c.bed_profile == LBM_BED_FLAT || c.bed_profile == LBM_BED_PIECEWISE;
            if (!(fixed && h->slot_valid[slot])) {
                launch_bed2d(c, h->b, (float)((double)(step + 1) * c.dt_phys), slot, bs);  // CODEX:3874
                h->slot_valid[slot] = 1;
            }
        }
        chain_waiting(h0->dev, bs, true, [&] { launch_wave_batch(storage, cm, (const char*)buf.dev + off[s] * E, cnt[s], blocks[s], bs); });
        for (int i = 0; i < n; i++) {
            lbm_handle* h = hs[i];
            if (s >= nsteps[i]) continue;
            h->sched.epoch++;
            h->cur ^= 1;
            h->post_stream = 0;
            h->step++;
        }
    }
    // every handle's later work comes after
    if (order) {
        if (!h0->ev_order) CHECK(hipEventCreateWithFlags(&h0->ev_order, hipEventDisableTiming));
        CHECK(hipEventRecord(h0->ev_order, bs));
        for (int i = 0; i < n; i++)
            if (hs[i]->stream != bs) CHECK(hipStreamWaitEvent(hs[i]->stream, h0->ev_order, 0));
    }
    return LBM_OK;
}

}  // namespace lbm

extern "C" {

int lbm_step_batch(lbm_handle_t* const* hs, const int32_t* nsteps, int32_t n) {
    if (!hs || !nsteps || n < 1 || !hs[0]) return LBM_E_ARG;
    lbm_handle* h0 = hs[0];
    API_BEGIN(h0)
    int rc = step_batch(hs, nsteps, n, h0->stream, h0->batch, true);
    if (rc) return rc;
    API_END(h0)
}

int lbm_upload_field(lbm_handle_t* h, int32_t field, const void* src, size_t bytes) {
    API_BEGIN(h)
    if (!src) return fail(h, LBM_E_ARG, "src is NULL");
    h->staged = StagedState{};  // any field set from outside: recompute what the staged step reuses
    if (field == LBM_FIELD_PIECEWISE_BED_Z)
        for (int sl = 0; sl < BED_RING; sl++) h->slot_valid[sl] = 0;
    const lbm_config_t& c = h->cfg;
    DevBufs& b = h->b;
    size_t n = b.n, n2 = b.n2, R = real_size(c);
    void* dst = nullptr;
    size_t need = 0;
    bool water_only = false;
    switch (field) {
        case LBM_FIELD_F: need = 19 * n * R; break;
        case LBM_FIELD_RHO: dst = b.rho; need = n * R; break;
        case LBM_FIELD_UX: dst = b.ux; need = n * R; break;
        case LBM_FIELD_UY: dst = b.uy; need = n * R; break;
        case LBM_FIELD_UZ: dst = b.uz; need = n * R; break;
        case LBM_FIELD_OMEGA: dst = b.omega; need = n * R; break;
        case LBM_FIELD_U_INLET_PROFILE: dst = b.u_inlet; need = c.nz * R; break;
        case LBM_FIELD_INFLOW_UX: dst = b.inflow; need = (size_t)c.nz * c.ny * R; break;
        case LBM_FIELD_PIECEWISE_BED_Z: dst = b.piecewise; need = c.nx * R; break;
        case LBM_FIELD_FS_FILL: dst = b.fill; need = n * R; water_only = true; break;
        case LBM_FIELD_FS_MASS: dst = b.mass; need = n * R; water_only = true; break;
        case LBM_FIELD_FS_TYPE: dst = b.type; need = n * 4; water_only = true; break;
        case LBM_FIELD_FS_TYPE_PREV: dst = b.type_prev; need = n * 4; water_only = true; break;
        case LBM_FIELD_FS_H: dst = b.fs_h; need = n2 * R; water_only = true; break;
        default: return fail(h, LBM_E_ARG, "field " + std::to_string(field) + " cannot be uploaded");
    }
    if (water_only && !c.fs_enabled) return fail(h, LBM_E_ARG, "free-surface fields exist only in water runs");
    if (bytes != need)
        return fail(h, LBM_E_ARG, "field " + std::to_string(field) + ": expected " + std::to_string(need) + " bytes, got " +
                                      std::to_string(bytes));
    if (field == LBM_FIELD_F) {
        if (c.storage == LBM_STORE_FP32 || c.storage == LBM_STORE_FP64) {
            CHECK(hipMemcpyAsync(b.f[h->cur], src, need, hipMemcpyHostToDevice, h->stream));
        } else {
            ensure_ftmp(h);
            CHECK(hipMemcpyAsync(b.ftmp, src, need, hipMemcpyHostToDevice, h->stream));
            launch_store_populations(c, b, h->cur, b.ftmp, h->step, h->stream);
        }
        h->post_stream = 1;
        h->initialized = 1;
    } else {
        CHECK(hipMemcpyAsync(dst, src, need, hipMemcpyHostToDevice, h->stream));
        if (field == LBM_FIELD_PIECEWISE_BED_Z && c.bed_profile == LBM_BED_PIECEWISE) {
            double mx = -1e300;
            for (int i = 0; i < c.nx; i++)
                mx = std::max(mx, R == 8 ? ((const double*)src)[i] : (double)((const float*)src)[i]);
            h->piecewise_max = mx;
            h->wall_top = wall_band_top(c, mx);
            // init_constants seeds the bed height with the profile (CODEX:1050-1054)
            std::vector<char> plane(n2 * R);
            for (size_t j = 0; j < (size_t)c.ny; j++) memcpy(plane.data() + j * c.nx * R, src, c.nx * R);
            for (int sl = 0; sl < b.ring; sl++)
                CHECK(hipMemcpyAsync(b.h[sl], plane.data(), n2 * R, hipMemcpyHostToDevice, h->stream));
        }
    }
    CHECK(hipStreamSynchronize(h->stream));
    API_END(h)
}

int lbm_download_field(lbm_handle_t* h, int32_t field, void* dst, size_t bytes) {
    API_BEGIN(h)
    if (!dst) return fail(h, LBM_E_ARG, "dst is NULL");
    const lbm_config_t& c = h->cfg;
    DevBufs& b = h->b;
    size_t n = b.n, n2 = b.n2, R = real_size(c);
    int slot = slot_of(h, h->step);
    const void* src = nullptr;
    size_t need = 0;
    std::vector<void*> tmp;
    auto scratch = [&](size_t sz) {
        void* p = nullptr;
        CHECK(hipMalloc(&p, sz));
        tmp.push_back(p);
        return p;
    };
    struct Cleanup {
        std::vector<void*>* v;
        ~Cleanup() {
            for (void* p : *v) (void)hipFree(p);
        }
    } cleanup{&tmp};
    bool water_only = false;
    switch (field) {
        case LBM_FIELD_F: need = 19 * n * R; break;
        case LBM_FIELD_RHO: src = b.rho; need = n * R; break;
        case LBM_FIELD_UX: src = b.ux; need = n * R; break;
        case LBM_FIELD_UY: src = b.uy; need = n * R; break;
        case LBM_FIELD_UZ: src = b.uz; need = n * R; break;
        case LBM_FIELD_OMEGA: src = b.omega; need = n * R; break;
        case LBM_FIELD_NU_T: src = b.nu_t; need = n * R; break;
        case LBM_FIELD_SMAG: src = b.smag; need = n * R; break;
        case LBM_FIELD_PHI:
        case LBM_FIELD_LATTICE_OPEN:
        case LBM_FIELD_LATTICE_OPEN_FRAC:
        case LBM_FIELD_WALL_TYPE:
        case LBM_FIELD_NEAR_OBSTACLE: {
            int32_t* lo = (int32_t*)scratch(19 * n * 4);
            void* lf = scratch(19 * n * R);
            int32_t* wt = (int32_t*)scratch(19 * n * 4);
            int32_t* nr = (int32_t*)scratch(n * 4);
            void* phi = scratch(n * R);
            launch_geometry_fields(c, b, slot, lo, lf, wt, nr, phi, h->stream);
            if (field == LBM_FIELD_PHI) src = phi, need = n * R;
            if (field == LBM_FIELD_LATTICE_OPEN) src = lo, need = 19 * n * 4;
            if (field == LBM_FIELD_LATTICE_OPEN_FRAC) src = lf, need = 19 * n * R;
            if (field == LBM_FIELD_WALL_TYPE) src = wt, need = 19 * n * 4;
            if (field == LBM_FIELD_NEAR_OBSTACLE) src = nr, need = n * 4;
            break;
        }
        case LBM_FIELD_OBSTACLE: {
            int* o = (int*)scratch(n * 4);
            launch_obstacle(c, b, slot, o, h->stream);
            src = o;
            need = n * 4;
            break;
        }
        case LBM_FIELD_WATER_H: src = b.h[slot]; need = n2 * R; break;
        case LBM_FIELD_U_WAVE: src = b.uw[slot]; need = n2 * R; break;
        case LBM_FIELD_V_WAVE: src = b.vw[slot]; need = n2 * R; break;
        case LBM_FIELD_W_WAVE: src = b.ww[slot]; need = n2 * R; break;
        case LBM_FIELD_U_INLET_PROFILE: src = b.u_inlet; need = c.nz * R; break;
        case LBM_FIELD_INFLOW_UX: src = b.inflow; need = (size_t)c.nz * c.ny * R; break;
        case LBM_FIELD_PIECEWISE_BED_Z: src = b.piecewise; need = c.nx * R; break;
        case LBM_FIELD_FS_FILL: src = b.fill; need = n * R; water_only = true; break;
        case LBM_FIELD_FS_TYPE: src = b.type; need = n * 4; water_only = true; break;
        case LBM_FIELD_FS_TYPE_PREV: src = b.type_prev; need = n * 4; water_only = true; break;
        case LBM_FIELD_FS_MASS: src = b.mass; need = n * R; water_only = true; break;
        case LBM_FIELD_FS_MASS_EXCESS: src = b.excess; need = n * R; water_only = true; break;
        case LBM_FIELD_FS_H: src = b.fs_h; need = n2 * R; water_only = true; break;
        case LBM_FIELD_VOF_EXTERIOR_GAS: src = b.ext_gas; need = n * 4; water_only = true; break;
        case LBM_FIELD_PRESSURE_REF_HEAD:
        case LBM_FIELD_GRAVITY_FRACTION: need = n * R; break;
        default: return fail(h, LBM_E_ARG, "unknown field " + std::to_string(field));
    }
    if (bytes != need)
        return fail(h, LBM_E_ARG, "field " + std::to_string(field) + ": expected " + std::to_string(need) + " bytes, got " +
                                      std::to_string(bytes));
    if (field == LBM_FIELD_NU_T || field == LBM_FIELD_SMAG) {
        if (!src) {  // the fused path keeps no LES diagnostics
            memset(dst, 0, need);
            return LBM_OK;
        }
    }
    if (field == LBM_FIELD_PRESSURE_REF_HEAD || field == LBM_FIELD_GRAVITY_FRACTION) {
        // update_pressure_reference_fields (CODEX:710-740) on the host from
        // the open mask and cell types
        std::vector<int> lo0(n, 1), ty(n, LBM_FS_LIQUID);
        int* o = (int*)scratch(n * 4);
        launch_obstacle(c, b, slot, o, h->stream);
        CHECK(hipMemcpyAsync(lo0.data(), o, n * 4, hipMemcpyDeviceToHost, h->stream));
        if (c.fs_enabled) CHECK(hipMemcpyAsync(ty.data(), b.type, n * 4, hipMemcpyDeviceToHost, h->stream));
        CHECK(hipStreamSynchronize(h->stream));
        for (size_t id = 0; id < n; id++) {
            size_t k = id / n2;
            float zc = ((float)k + 0.5f) * c.dx;
            double head = 0.0, gf = 0.0;
            bool open = lo0[id] == 0;  // obstacle mask: 0 = open
            if (c.fs_enabled && open && ty[id] != LBM_FS_GAS) {
                if (c.pressure_formulation == LBM_PF_HYDROSTATIC_BALANCED) {
                    head = (float)(c.free_surface_base_m - zc);
                } else if (c.pressure_formulation == LBM_PF_VOF_COMPONENT_BALANCED) {
                    if (zc < c.free_surface_base_m)
                        head = (float)(c.free_surface_base_m - zc);
                    else
                        gf = 1.0;
                } else {
                    gf = 1.0;
                }
            }
            double v = field == LBM_FIELD_PRESSURE_REF_HEAD ? head : gf;
            if (R == 8)
                ((double*)dst)[id] = v;
            else
                ((float*)dst)[id] = (float)v;
        }
        return LBM_OK;
    }
    if (water_only && !c.fs_enabled) return fail(h, LBM_E_ARG, "free-surface fields exist only in water runs");
    if (field == LBM_FIELD_F) src = post_stream_populations(h);
    CHECK(hipMemcpyAsync(dst, src, need, hipMemcpyDeviceToHost, h->stream));
    CHECK(hipStreamSynchronize(h->stream));
    check_wave_flag(h);
    API_END(h)
}

int lbm_get_counters(lbm_handle_t* h, lbm_counters_t* out) {
    API_BEGIN(h)
    if (!out) return fail(h, LBM_E_ARG, "out is NULL");
    if (h->path == LBM_PATH_FUSED && h->step > 0) launch_cut_link_count(h->cfg, h->b, slot_of(h, h->step), h->stream);
    int ic[IC_COUNT];
    double dc[DC_COUNT];
    CHECK(hipMemcpyAsync(ic, h->b.icnt, sizeof(ic), hipMemcpyDeviceToHost, h->stream));
    CHECK(hipMemcpyAsync(dc, h->b.dcnt, sizeof(dc), hipMemcpyDeviceToHost, h->stream));
    CHECK(hipStreamSynchronize(h->stream));
    memset(out, 0, sizeof(*out));
    out->boundary_cut_link_count = ic[IC_BOUNDARY_CUT];
    out->boundary_refill_count = ic[IC_BOUNDARY_REFILL];
    out->free_surface_refill_count = ic[IC_FS_REFILL];
    out->free_surface_coalescence_repair_count = ic[IC_FS_REPAIR];
    out->free_surface_thin_gap_bridge_count = ic[IC_FS_BRIDGE];
    out->free_surface_cut_link_count = ic[IC_FS_CUT];
    out->free_surface_vof_excess_count = ic[IC_VOF_EXCESS];
    out->free_surface_vof_orphan_count = ic[IC_VOF_ORPHAN];
    out->free_surface_vof_collapse_cell_count = ic[IC_VOF_COLLAPSE_CELLS];
    out->free_surface_vof_detached_advect_count = ic[IC_VOF_ADVECT];
    out->free_surface_vof_detached_compress_count = ic[IC_VOF_COMPRESS];
    out->pad0 = ic[IC_FLOOD_UNCONVERGED];
    out->free_surface_vof_collapse_candidate_volume_m3 = dc[DC_CAND_VOLUME];
    out->free_surface_vof_collapse_applied_volume_m3 = dc[DC_APPLIED_VOLUME];
    out->free_surface_vof_collapse_removal_scale = dc[DC_REMOVAL_SCALE];
    out->free_surface_vof_collapse_candidate_mass = dc[DC_CAND_MASS];
    out->free_surface_vof_collapse_exterior_capacity = dc[DC_EXT_CAPACITY];
    out->free_surface_vof_detached_advect_mass = dc[DC_ADVECT_MASS];
    out->free_surface_vof_detached_compress_mass = dc[DC_COMPRESS_MASS];
    API_END(h)
}

static int reduce_checks(lbm_handle* h, int32_t what, const double* out) {
    if (!out) return fail(h, LBM_E_ARG, "out is NULL");
    if (what < 0 || what > LBM_REDUCE_NONFINITE) return fail(h, LBM_E_ARG, "unknown reduction");
    if ((what == LBM_REDUCE_VOF_MASS || what == LBM_REDUCE_VOF_VOLUME) && !h->cfg.fs_enabled)
        return fail(h, LBM_E_ARG, "VOF reductions need a water run");
    return LBM_OK;
}

// the reduction into the device slot dst; part: partial sums (max_parts)
static void reduce_into(lbm_handle* h, int32_t what, double* part, size_t max_parts, double* dst, double* out) {
    launch_cellflags(h->cfg, h->b, slot_of(h, h->step), h->stream);
    launch_reduce(h->cfg, h->b, h->cur, h->post_stream, slot_of(h, h->step), what, part, max_parts, dst, h->stream);
    double v = 0.0;
    CHECK(hipMemcpyAsync(&v, dst, sizeof(double), hipMemcpyDeviceToHost, h->stream));
    CHECK(hipStreamSynchronize(h->stream));
    check_wave_flag(h);
    *out = v;
}

int lbm_reduce(lbm_handle_t* h, int32_t what, double* out) {
    API_BEGIN(h)
    if (int e = reduce_checks(h, what, out)) return e;
    reduce_into(h, what, h->b.rpart, 3 * reduce_chunks(h->b.n), h->b.dcnt + DC_REDUCE, out);
    API_END(h)
}

int lbm_reduce_noalloc(lbm_handle_t* h, int32_t what, double* out) {
    API_BEGIN(h)
    if (int e = reduce_checks(h, what, out)) return e;
    if (h->b.n > LBM_REDUCE_NOALLOC_MAX_CELLS || reduce_chunks(h->b.n) > h->noalloc_parts)
        return fail(h, LBM_E_ARG, "lbm_reduce_noalloc: the grid has " + std::to_string(h->b.n) + " cells, more than the " +
                                      std::to_string((size_t)LBM_REDUCE_NOALLOC_MAX_CELLS) +
                                      " its static scratch serves; use lbm_reduce");
    // partial sums in the static array, the result in the slot after them
    reduce_into(h, what, h->noalloc_part, h->noalloc_parts, h->noalloc_part + h->noalloc_parts, out);
    API_END(h)
}

int lbm_get_step(lbm_handle_t* h, int64_t* out) {
    if (!h || !out) return LBM_E_ARG;
    *out = h->step;
    return LBM_OK;
}

int lbm_sync(lbm_handle_t* h) {
    API_BEGIN(h)
    CHECK(hipStreamSynchronize(h->stream));
    check_wave_flag(h);
    API_END(h)
}

int lbm_query(lbm_handle_t* h, int32_t* idle) {
    if (!h || !idle) return LBM_E_ARG;
    hipError_t e = hipStreamQuery(h->stream);
    if (e == hipSuccess) {
        *idle = 1;
        return LBM_OK;
    }
    if (e == hipErrorNotReady) {
        *idle = 0;
        return LBM_OK;
    }
    return fail(h, LBM_E_HIP, std::string("hipStreamQuery: ") + hipGetErrorString(e));
}

int lbm_last_step_timing(lbm_handle_t* h, double* ms_per_step, double* ms_main_kernel) {
    if (!h) return LBM_E_ARG;
    if (!h->timing) return fail(h, LBM_E_STATE, "timing is off (lbm_set_option(h, \"timing\", 1))");
    if (ms_per_step) *ms_per_step = h->last_ms_step;
    if (ms_main_kernel) *ms_main_kernel = h->last_ms_kernel;
    return LBM_OK;
}

int lbm_device_bytes(const lbm_config_t* cfg, size_t* out) {
    const char* why = validate(cfg);
    if (why || !out) return LBM_E_CONFIG;
    *out = device_bytes(*cfg, resolve_path(*cfg));
    return LBM_OK;
}

int lbm_set_option(lbm_handle_t* h, const char* key, int64_t value) {
    if (!h || !key) return LBM_E_ARG;
    if (!strcmp(key, "timing")) {
        h->timing = value ? 1 : 0;
        return LBM_OK;
    }
    if (!strcmp(key, "graph")) {
        h->graph = value ? 1 : 0;
        return LBM_OK;
    }
    if (!strcmp(key, "wavefront")) {
        h->sched.wavefront = value ? 1 : 0;
        return LBM_OK;
    }
    if (!strcmp(key, "resident")) {
        if (value < 0 || value > 2) return fail(h, LBM_E_ARG, "resident must be 0, 1 or 2 (chosen by timing)");
        h->sched.resident = value == 1 ? 1 : 0;
        h->sched_auto = value == 2 ? 1 : 0;
        return LBM_OK;
    }
    if (!strcmp(key, "wave_lag")) {
        if (value < 0 || value > 4096) return fail(h, LBM_E_ARG, "wave_lag must be in [0, 4096] (0: from the tiles a device holds)");
        h->sched.lag = (int)value;
        return LBM_OK;
    }
    if (!strcmp(key, "vof_band")) {
        h->vof_band = value ? 1 : 0;
        return LBM_OK;
    }
    if (!strcmp(key, "flood_sync")) {
        h->flood_sync = value ? 1 : 0;
        return LBM_OK;
    }
    if (!strcmp(key, "xcd_strips")) {
        h->sched.strips = value ? 1 : 0;
        return LBM_OK;
    }
    if (!strcmp(key, "block_x")) {
        if (value != 64 && value != 128 && value != 256) return fail(h, LBM_E_ARG, "block_x must be 64, 128 or 256");
        h->block_x = (int)value;
        return LBM_OK;
    }
    return fail(h, LBM_E_ARG, std::string("unknown option ") + key);
}

int lbm_get_option(lbm_handle_t* h, const char* key, int64_t* out) {
    if (!h || !key || !out) return LBM_E_ARG;
    if (!strcmp(key, "resident")) *out = h->sched_auto ? 2 : h->sched.resident;
    else if (!strcmp(key, "timing")) *out = h->timing;
    else if (!strcmp(key, "graph")) *out = h->graph;
    else if (!strcmp(key, "wavefront")) *out = h->sched.wavefront;
    else if (!strcmp(key, "wave_lag")) *out = h->sched.lag;
    else if (!strcmp(key, "xcd_strips")) *out = h->sched.strips;
    else if (!strcmp(key, "block_x")) *out = h->block_x;
    else if (!strcmp(key, "vof_band")) *out = h->vof_band;
    else if (!strcmp(key, "flood_sync")) *out = h->flood_sync;
    else return fail(h, LBM_E_ARG, std::string("unknown option ") + key);
    return LBM_OK;
}

int lbm_get_path(lbm_handle_t* h, int32_t* out) {
    if (!h || !out) return LBM_E_ARG;
    *out = h->path;
    return LBM_OK;
}

}  // extern "C"
