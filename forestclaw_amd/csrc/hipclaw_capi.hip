// C ABI of the batched patch-update backend (declared in include/hipclaw.h).
//
// Device memory model: one allocation per slab, sized once for max_patches.
// q lives in two slabs (ping-pong) unless a single tile covers a patch, in
// which case the step is done in place.  All launches go to one
// non-blocking stream of default priority owned by the state.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cstring>
#include <cmath>
#include <mutex>
#include <unordered_map>
#include <vector>

#include "../../include/hipclaw.h"
#include "../../include/hipclaw_solver.h"
#include "hipclaw_internal.h"

using hipclaw::GhostOp;

namespace {

thread_local char g_err[512] = "";

int fail(int code, const char* fmt, const char* detail = "") {
    std::snprintf(g_err, sizeof(g_err), fmt, detail);
    return code;
}

int hipfail(hipError_t e, const char* where) {
    std::snprintf(g_err, sizeof(g_err), "%s: %s", where, hipGetErrorString(e));
    return HIPCLAW_EDEVICE;
}

#define HIP_TRY(call)                                  \
    do {                                               \
        hipError_t _e = (call);                        \
        if (_e != hipSuccess) return hipfail(_e, #call); \
    } while (0)

}  // namespace

// One kernel launch of a ghost-fill program: exchange ops or physical-BC ops
// start .. start+nops-1 of the program's device array, or per-destination
// groups group0 .. group0+ngroups-1 of its group array.
enum { GHOST_LAUNCH_EXCHANGE, GHOST_LAUNCH_PHYSBC, GHOST_LAUNCH_GROUPS };
struct GhostLaunch {
    int start = 0, nops = 0;
    int kind = GHOST_LAUNCH_EXCHANGE;
    int group0 = 0, ngroups = 0;
};

// One ghost-fill program (hipclaw_set_ghost_program_k).
struct GhostProgram {
    GhostOp* ops = nullptr;
    int nops = 0, cap = 0;
    int4* groups = nullptr;
    int groups_cap = 0;
    // two layouts of the same program (hipclaw_set_ghost_layout): 0 = one
    // launch per phase, 1 = per-destination groups
    GhostLaunch launch[2][6];
    int nlaunch[2] = {0, 0};
    bool layouts_built = false;
    bool set = false;             // installed since the last regrid
};

// One time-synchronisation program (hipclaw_set_sync_program_k).
struct SyncProgram {
    int* ops = nullptr;       // [nops][8], grouped by destination
    int* start = nullptr;     // [ngroups + 1]
    int* resets = nullptr;    // [nresets][2] (slot, face)
    int ngroups = 0, nresets = 0, ops_cap = 0, groups_cap = 0, resets_cap = 0;
    int maxgroup = 0;         // most ops of one destination
    bool set = false;         // installed since the last regrid
};

struct hipclaw_state {
    hipclaw_config_t cfg;
    int tile = 16, ntx = 1, nty = 1;
    bool inplace = false;
    bool conservative = false;   // the solver has the FP32-bulk flux form
    int max_patches = 0, npatch = 0;
    long long pstride = 0, astride = 0;
    double* q[2] = {nullptr, nullptr};
    int cur = 0;
    double* aux = nullptr;
    double* dxs = nullptr;
    double* dys = nullptr;
    int* all_slots = nullptr;
    int* batch_slots = nullptr;
    GhostProgram gp[HIPCLAW_MAX_PROGRAMS];
    unsigned long long* cfl = nullptr;
    double* mass_buf = nullptr;
    // time synchronisation: flux registers and the per-destination op groups
    double* reg = nullptr;
    long long rstride = 0;
    SyncProgram sp[HIPCLAW_MAX_PROGRAMS];
    bool reg_accumulate = false;  // subcycling: registers add up over substeps
    // subcycling: the state before each patch's last update, and time-interpolated copies
    double* qlast = nullptr;
    double* qti = nullptr;
    double* dtscale = nullptr;    // per-slot factor on dt, or null
    int ghost_layout = 0;         // launch layout of the ghost fill (hipclaw_set_ghost_layout)
    int waves = 0;                // waves per workgroup of 16 x 16 tiles, 0 = the default
    // subcycled advance: slots grouped by level (hipclaw_set_levels); level
    // lmin + l holds lslots[loff[l] .. loff[l + 1])
    int* lslots = nullptr;
    int lslots_cap = 0;
    std::vector<int> loff;
    int lmin = 0, lmax = -1;      // lmax < lmin: no levels set
    void* hstage = nullptr;       // pinned staging buffer of host-to-device uploads
    size_t hstage_cap = 0;
    bool hstage_busy = false;     // a copy out of hstage may still run
    double* qsave = nullptr;      // saved state (hipclaw_save_step), max_patches slabs
    bool saved = false;           // a saved state of the current slots exists
    // With two slabs the saved state stays in slab save_slab, uncopied, while
    // nothing writes that slab (a one-step advance writes the other one);
    // anything else that may write it copies it into qsave first.
    bool save_lazy = false;
    int save_slab = 0;
    // ghost-patch message: slab offsets of one field's band, in message order
    int* band = nullptr;
    int nband = 0;
    hipStream_t st = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int graph_steps = 0;
    int graph_cur = 0;            // the slab a captured advance graph starts from
    // timing of the update kernel inside advance_async
    bool time_kernels = true;
    std::vector<hipEvent_t> kev;  // pairs (kernel start, kernel end) per step
    int timed_steps = 0;
    hipEvent_t region[2] = {nullptr, nullptr};
    // host copies of program 0 (ghost fill with its phases, time sync): the
    // fused step is built from them
    std::vector<int> gp0_ops, sp0_ops;
    // the fused step's sync ops, FP32 flux form (same-size ops paired);
    // other forms run sync program 0 as it is
    int* fsops = nullptr;
    int fsops_cap = 0, nfsync = 0;
    bool fpaired = false;
    int gp0_phase[7] = {0, 0, 0, 0, 0, 0, 0};
    bool gp0_set = false;
    // fused step (ghost layout 2): the update pulls same-size ghost strips, one
    // ghost_flat launch fills the rest and computes the time-sync increments,
    // which the next update adds (deferred through qcorr)
    bool fused_ok = false;
    // ops: corners of every step (averages, interpolations) [0, fc0), the
    // other ops of every step [fc0, fc0 + fbig), same-size face copies
    // [.., + fcopy), corner copies [.., + fc1) (copies only in an advance's
    // last step)
    GhostOp* fops = nullptr;
    int fops_cap = 0, fc0 = 0, fbig = 0, fcopy = 0, fc1 = 0;
    int* pull = nullptr;          // [slot][8] offsets of same-size ghost strips (StepArgs::pull)
    int pull_cap = 0;
    bool pull_ok = false;
    hipclaw::FaceRel* frel = nullptr;
    int frel_cap = 0;
    int* cmask = nullptr;
    int cmask_cap = 0;
    int* qslots = nullptr;        // the patches with a corrected face (apply_qcorr)
    int qslots_cap = 0, nqslots = 0;
    double* qcorr = nullptr;
    long long cstride = 0;
    int qcorr_cap = 0;            // patches qcorr holds
    bool fsync = false;           // the fused step applies time-sync corrections
    // device ints for the regrid's tags and transfer ops, grown as needed and
    // kept (a regrid allocates nothing when the forest fits)
    int* iwork = nullptr;
    size_t iwork_cap = 0;
};

// ---------------------------------------------------------------- allocation-free step
// Device scratch of hipclaw_step_batch_noalloc: the CFL word, then the staged
// slot list.  Every GPU holds its own copy of this array for the life of the
// process; its device address is looked up once per device (scratch_ptr) and
// only that address is passed to kernels, copies and fills.
namespace {
constexpr size_t NOALLOC_CFL_BYTES = 256;
constexpr size_t NOALLOC_SCRATCH_BYTES =
    NOALLOC_CFL_BYTES + sizeof(int) * (size_t) HIPCLAW_NOALLOC_MAX_SLOTS;
constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// The device address of `scratch` on the current device, looked up on the
// first call per device; nullptr (and the error text set) on failure.
static void* scratch_ptr(int* rc) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) { *rc = hipfail(e, "hipGetDevice"); return nullptr; }
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) {
        *rc = fail(HIPCLAW_ENODEV, "device index beyond the scratch table%s");
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *rc = hipfail(e != hipSuccess ? e : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)");
            return nullptr;
        }
        g_scratch_ptr[dev] = p;
    }
    *rc = HIPCLAW_OK;
    return g_scratch_ptr[dev];
}

// Program 0's fused tables are built in two passes over its ops: the first
// checks every op and fills the face table in the staging buffer, the second
// (once the state has taken the program) builds the flat op list and the pull
// offsets beside it.
struct FusedBuild {
    bool fusable = false;
    int nlite = 0;              // ops of the flat list before the copies
    int ncorner = 0;            // of them: average and interpolate corners
    int ncopy = 0, ncopyc = 0;  // face and corner copies
    std::vector<int> bcop;      // per patch: the physical-BC op kept (the last), -1 if none
    char* stage = nullptr;      // face table, flat op list, pull offsets
    size_t off_fops = 0, off_pull = 0;
};
static int fused_check_and_faces(hipclaw_state_t* s, const int* ops, int nops, const int* phase_start,
                                 FusedBuild& b);
static int fused_finish(hipclaw_state_t* s, const int* ops, int nops, FusedBuild& b);
static int build_fused_sync(hipclaw_state_t* s);

// Host-to-device copies of programs and tables go through the state's pinned
// staging buffer (pageable copies run several times slower).  The buffer is
// reused by the next upload, which first waits for the copies out of it.

// The staging buffer with room for `bytes`, once no copy out of it runs any
// more (the caller may then fill it and copy out of it).
static int stage_reserve(hipclaw_state_t* s, size_t bytes, char** base) {
    if (s->hstage_busy) {
        HIP_TRY(hipStreamSynchronize(s->st));
        s->hstage_busy = false;
    }
    if (bytes > s->hstage_cap) {
        if (s->hstage) hipHostFree(s->hstage);
        s->hstage = nullptr;
        s->hstage_cap = 0;
        const size_t cap = bytes + bytes / 4;
        if (hipHostMalloc(&s->hstage, cap, hipHostMallocDefault) != hipSuccess)
            return fail(HIPCLAW_ENOMEM, "hipHostMalloc of the upload staging buffer failed%s");
        s->hstage_cap = cap;
    }
    *base = (char*) s->hstage;
    return HIPCLAW_OK;
}

// Room for n elements plus a quarter: the tables of a growing forest are not
// reallocated (hipFree waits for the device) at every regrid.
static int grown_cap(int n) {
    const long long c = (long long) n + n / 4 + 16;
    return c > INT_MAX ? INT_MAX : (int) c;
}

// A device array of at least n elements (grown, contents dropped, as needed).
template <typename T>
static int ensure_dev(T*& dev, int& cap, int n, const char* what) {
    if (n > cap) {
        hipFree(dev);
        dev = nullptr;
        cap = 0;
        const int c = grown_cap(n);
        if (hipMalloc((void**) &dev, sizeof(T) * (size_t) c) != hipSuccess)
            return fail(HIPCLAW_ENOMEM, "hipMalloc of %s failed", what);
        cap = c;
    }
    return HIPCLAW_OK;
}

// Copy n elements out of the staging buffer (at `host`, inside it) into a
// device array grown as needed; the copy is asynchronous on the state's stream.
template <typename T>
static int stage_copy(hipclaw_state_t* s, T*& dev, int& cap, const T* host, int n, const char* what) {
    if (int rc = ensure_dev(dev, cap, n, what)) return rc;
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(dev, host, sizeof(T) * (size_t) n, hipMemcpyHostToDevice, s->st));
        s->hstage_busy = true;
    }
    return HIPCLAW_OK;
}

// Copy `bytes` from host memory to `dev` through the staging buffer.
static int staged_upload(hipclaw_state_t* s, void* dev, const void* host, size_t bytes) {
    char* base = nullptr;
    if (int rc = stage_reserve(s, bytes, &base)) return rc;
    std::memcpy(base, host, bytes);
    HIP_TRY(hipMemcpyAsync(dev, base, bytes, hipMemcpyHostToDevice, s->st));
    s->hstage_busy = true;
    return HIPCLAW_OK;
}

// Upload n elements of `host` into a device array grown as needed.
template <typename T>
static int upload_array(hipclaw_state_t* s, T*& dev, int& cap, const T* host, int n, const char* what) {
    if (int rc = ensure_dev(dev, cap, n, what)) return rc;
    if (n > 0) return staged_upload(s, dev, host, sizeof(T) * (size_t) n);
    return HIPCLAW_OK;
}

// Copy slab `from` into qsave (allocated on first use).
static int copy_to_save(hipclaw_state_t* s, int from) {
    const size_t bytes = sizeof(double) * (size_t) s->pstride * (size_t) s->npatch;
    if (!s->qsave) {
        if (hipMalloc((void**) &s->qsave, sizeof(double) * (size_t) s->pstride * (size_t) s->max_patches) != hipSuccess)
            return fail(HIPCLAW_ENOMEM, "hipMalloc of the saved state failed%s");
    }
    if (bytes > 0) HIP_TRY(hipMemcpyAsync(s->qsave, s->q[from], bytes, hipMemcpyDeviceToDevice, s->st));
    return HIPCLAW_OK;
}

// Before anything that may write slab save_slab: a lazily saved state is
// copied into qsave.  A one-step advance from the saved slab writes only the
// other one (the update's destination, then that slab's ghost cells and
// fix-up), so it leaves the saved state in place.
static int guard_save(hipclaw_state_t* s, bool one_step_advance = false) {
    if (!s->save_lazy) return HIPCLAW_OK;
    if (one_step_advance && s->cur == s->save_slab) return HIPCLAW_OK;
    s->save_lazy = false;
    return copy_to_save(s, s->save_slab);
}

// A captured graph holds its launches' arguments: op counts, table and
// buffer addresses, patch counts.  Whatever changes those drops the graph
// (replaying it is then an error until the next capture).
static void drop_graph(hipclaw_state_t* s) {
    if (s->gexec) { hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { hipGraphDestroy(s->graph); s->graph = nullptr; }
    s->graph_steps = 0;
}

extern "C" {

const char* hipclaw_last_error(void) { return g_err; }

size_t hipclaw_config_size(void) { return sizeof(hipclaw_config_t); }

int hipclaw_initialize_device(int device, char* name, size_t name_len) {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(HIPCLAW_ENODEV, "no HIP device visible%s");
    if (device < 0 || device >= count) return fail(HIPCLAW_EINVAL, "device index out of range%s");
    HIP_TRY(hipSetDevice(device));
    int rc;
    if (!scratch_ptr(&rc)) return rc;
    if (name && name_len > 0) {
        hipDeviceProp_t p;
        HIP_TRY(hipGetDeviceProperties(&p, device));
        std::snprintf(name, name_len, "%s (%s)", p.name, p.gcnArchName);
    }
    return HIPCLAW_OK;
}

// ---- Riemann solvers: the three built in (numbers 0..2), then those an
// application registers (HIPCLAW_RP_USER on).  Registration publishes the
// descriptor before the count, so readers need no lock.
static std::mutex g_solvers_mu;
static const hipclaw_solver_t* g_solvers[HIPCLAW_MAX_USER_SOLVERS];
static std::atomic<int> g_nsolvers{0};

static const hipclaw_solver_t* user_solver(int id) {
    const int k = id - HIPCLAW_RP_USER;
    if (k < 0 || k >= g_nsolvers.load(std::memory_order_acquire)) return nullptr;
    return g_solvers[k];
}

int hipclaw_register_solver(const hipclaw_solver_t* d, int* id) {
    if (!d || !id) return fail(HIPCLAW_EINVAL, "hipclaw_register_solver: bad arguments%s");
    if (d->abi != HIPCLAW_SOLVER_ABI || d->step_args_size != (int) sizeof(hipclaw::StepArgs))
        return fail(HIPCLAW_EINVAL, "hipclaw_register_solver: solver built against another "
                                    "version of the step kernel%s");
    if (d->meqn < 1 || d->mwaves < 1 || d->mwaves > 10 || d->maux < 0 || !d->launch_fp64 ||
        !(d->tiles & (8 | 16)))
        return fail(HIPCLAW_EINVAL, "hipclaw_register_solver: bad descriptor%s");
    std::lock_guard<std::mutex> lock(g_solvers_mu);
    const int n = g_nsolvers.load(std::memory_order_relaxed);
    for (int k = 0; k < n; k++)
        if (g_solvers[k] == d) { *id = HIPCLAW_RP_USER + k; return HIPCLAW_OK; }
    if (n >= HIPCLAW_MAX_USER_SOLVERS) return fail(HIPCLAW_EINVAL, "hipclaw_register_solver: registry full%s");
    g_solvers[n] = d;
    g_nsolvers.store(n + 1, std::memory_order_release);
    *id = HIPCLAW_RP_USER + n;
    return HIPCLAW_OK;
}

const hipclaw_solver_t* hipclaw_registered_solver(int id) { return user_solver(id); }

namespace {
struct SolverShape {
    int meqn, mwaves, maux;
    bool conservative;   // FP32-bulk flux form
    bool time_sync;      // keeps time-sync registers
};
}  // namespace

static bool solver_shape(int solver, SolverShape* sh) {
    static const SolverShape builtin[3] = {
        {1, 1, 2, false, true}, {3, 3, 0, true, true}, {4, 4, 0, true, true}};
    if (solver >= 0 && solver <= 2) { *sh = builtin[solver]; return true; }
    const hipclaw_solver_t* d = user_solver(solver);
    if (!d) return false;
    *sh = {d->meqn, d->mwaves, d->maux, d->conservative != 0, d->time_sync != 0};
    return true;
}

// Tile sizes instantiated for a solver in a precision.
static bool solver_tile_ok(int solver, int precision, int tile) {
    if (solver >= 0 && solver <= 2) return hipclaw::step_supported(solver, precision, tile);
    const hipclaw_solver_t* d = user_solver(solver);
    if (!d || (tile != 8 && tile != 16) || !(d->tiles & tile)) return false;
    return precision == HIPCLAW_FP64 ? d->launch_fp64 != nullptr : d->launch_fp32 != nullptr;
}

static int solver_default_waves(int solver, int precision, int tile) {
    if (solver >= 0 && solver <= 2)
        return precision == HIPCLAW_FP64 ? hipclaw::default_waves_fp64(solver, tile)
                                         : hipclaw::default_waves_fp32(solver, tile);
    return tile != 16 ? 0 : precision == HIPCLAW_FP64 ? hipclaw::default_waves_t16<double>()
                                                      : hipclaw::default_waves_t16<float>();
}

static int launch_solver_step(int solver, int precision, int tile, int waves,
                              const hipclaw::StepArgs& a, const hipclaw::StepLaunch& L) {
    if (solver >= 0 && solver <= 2)
        return precision == HIPCLAW_FP64 ? hipclaw::launch_step_fp64(solver, tile, waves, a, L)
                                         : hipclaw::launch_step_fp32(solver, tile, waves, a, L);
    const hipclaw_solver_t* d = user_solver(solver);
    const hipclaw_step_launcher_t fn = !d ? nullptr : precision == HIPCLAW_FP64 ? d->launch_fp64 : d->launch_fp32;
    if (!fn) return (int) hipErrorInvalidValue;
    return fn(tile, waves, &a, (void*) L.st, (void*) L.start, (void*) L.stop);
}

// The solver-dependent part of the step arguments.
static void set_method(hipclaw::StepArgs& a, const hipclaw_config_t* c) {
    a.order2 = c->order[0] == 2;
    a.transverse = c->order[1];
    for (int k = 0; k < 10; k++) a.mthlim[k] = k < c->mwaves ? c->mthlim[k] : 0;
    a.pp.grav = c->grav;
    a.pp.gamma = c->gamma;
    for (int k = 0; k < 8; k++) a.pp.p[k] = c->rp_params[k];
}

// Tiles per patch in one direction: the last tile of a patch whose size the
// tile does not divide overlaps the one before it (the step kernel moves it
// back to end at the patch edge).
static int tiles_along(int n, int t) { return (n + t - 1) / t; }

// Tile size measured fastest on MI355X for the update kernel (the tile sweep of
// the benchmark forest, mx = 16 and 32): 16 x 16 for both forms (four waves
// per workgroup for FP64, two for the FP32-bulk form); smaller patches take
// 8 x 8.  Where neither divides the patch, the tile with the fewest Riemann
// solves per patch, (T + 2) (T + 3) per tile, counting the overlapped tiles
// (ties to the larger).  hipclaw_autotune re-measures on the caller's patches.
static int default_tile(const hipclaw_config_t* c) {
    int best = 0;
    long long best_cost = 0;
    for (int t = 16; t >= 8; t >>= 1) {
        if (t > c->mx || t > c->my || !solver_tile_ok(c->solver, c->precision, t)) continue;
        const long long cost = (long long) tiles_along(c->mx, t) * tiles_along(c->my, t) * (t + 2) * (t + 3);
        if (!best || cost < best_cost) { best = t; best_cost = cost; }
    }
    return best ? best : 8;
}

// Checks a configuration for the step kernels; *tile_out = the tile to use.
static int check_config(const hipclaw_config_t* cfg, int* tile_out) {
    const int solver = cfg->solver;
    SolverShape sh;
    if (!solver_shape(solver, &sh)) return fail(HIPCLAW_EINVAL, "unknown Riemann solver%s");
    if (cfg->meqn != sh.meqn || cfg->mwaves != sh.mwaves || cfg->maux != sh.maux)
        return fail(HIPCLAW_EINVAL, "meqn/mwaves/maux do not match the Riemann solver%s");
    if (cfg->time_sync && !sh.time_sync)
        return fail(HIPCLAW_EINVAL, "the Riemann solver provides no conservative flux for time sync%s");
    if (cfg->mbc != 2) return fail(HIPCLAW_EINVAL, "the batched step requires mbc = 2%s");
    if (cfg->precision != HIPCLAW_FP64 && cfg->precision != HIPCLAW_FP32_BULK)
        return fail(HIPCLAW_EINVAL, "unknown precision%s");
    int tile = cfg->tile > 0 ? cfg->tile : default_tile(cfg);
    if (!solver_tile_ok(solver, cfg->precision, tile))
        return fail(HIPCLAW_EINVAL, "tile size not available for this solver and precision%s");
    if (cfg->time_sync && cfg->precision == HIPCLAW_FP64 && tile == 32)
        return fail(HIPCLAW_EINVAL, "FP64 tiles of 32 leave no LDS for the time-sync edge buffers%s");
    if (cfg->mx < tile || cfg->my < tile)
        return fail(HIPCLAW_EINVAL, "mx and my must be at least the tile size (8)%s");
    if (cfg->order[0] < 1 || cfg->order[0] > 2 || cfg->order[1] < 0 || cfg->order[1] > 2)
        return fail(HIPCLAW_EINVAL, "order must be (1|2, 0|1|2)%s");
    for (int k = 0; k < cfg->mwaves; k++)
        if (cfg->mthlim[k] < 0 || cfg->mthlim[k] > 6) return fail(HIPCLAW_EINVAL, "mthlim out of range%s");
    if (cfg->time_sync && cfg->mx != cfg->my)
        return fail(HIPCLAW_EINVAL, "time sync requires mx = my%s");
    *tile_out = tile;
    return HIPCLAW_OK;
}

int hipclaw_create(const hipclaw_config_t* cfg, int max_patches, hipclaw_state_t** out) {
    if (!cfg || !out || max_patches <= 0) return fail(HIPCLAW_EINVAL, "hipclaw_create: bad arguments%s");
    *out = nullptr;
    int tile = 0;
    if (int rc = check_config(cfg, &tile)) return rc;

    hipclaw_state* s = new hipclaw_state();
    s->cfg = *cfg;
    s->tile = tile;
    {
        SolverShape sh;
        solver_shape(cfg->solver, &sh);
        s->conservative = sh.conservative;
    }
    s->ntx = tiles_along(cfg->mx, tile);
    s->nty = tiles_along(cfg->my, tile);
    s->inplace = (s->ntx == 1 && s->nty == 1);
    s->max_patches = max_patches;
    // measured on MI355X: per-destination groups pay off once a patch's ghost
    // strips are long (mx >= 32); shorter ones keep one launch per phase
    s->ghost_layout = 2;
    const long long cells = (long long) (cfg->mx + 2 * cfg->mbc) * (cfg->my + 2 * cfg->mbc);
    s->pstride = cells * cfg->meqn;
    s->astride = cells * cfg->maux;
    auto cleanup = [&](int rc) { hipclaw_destroy(s); return rc; };
#define ALLOC(p, bytes)                                                         \
    do {                                                                        \
        if (hipMalloc((void**) &(p), (bytes)) != hipSuccess) {                  \
            return cleanup(fail(HIPCLAW_ENOMEM, "hipMalloc of %s failed", #p)); \
        }                                                                       \
    } while (0)
    if (hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(HIPCLAW_EDEVICE, "hipStreamCreate failed%s"));
    ALLOC(s->q[0], sizeof(double) * s->pstride * max_patches);
    if (!s->inplace) ALLOC(s->q[1], sizeof(double) * s->pstride * max_patches);
    if (cfg->maux > 0) ALLOC(s->aux, sizeof(double) * s->astride * max_patches);
    ALLOC(s->dxs, sizeof(double) * max_patches);
    ALLOC(s->dys, sizeof(double) * max_patches);
    ALLOC(s->all_slots, sizeof(int) * max_patches);
    ALLOC(s->batch_slots, sizeof(int) * max_patches);
    ALLOC(s->cfl, sizeof(unsigned long long));
    ALLOC(s->mass_buf, sizeof(double) * 2 * cfg->meqn * (size_t) max_patches);
    if (cfg->subcycle) {
        ALLOC(s->qlast, sizeof(double) * s->pstride * max_patches);
        ALLOC(s->qti, sizeof(double) * s->pstride * max_patches);
        if (hipMemset(s->qlast, 0, sizeof(double) * s->pstride * max_patches) != hipSuccess ||
            hipMemset(s->qti, 0, sizeof(double) * s->pstride * max_patches) != hipSuccess)
            return cleanup(fail(HIPCLAW_EDEVICE, "initial device copies failed%s"));
        s->reg_accumulate = true;
    }
    if (cfg->time_sync) {
        s->rstride = 16LL * cfg->mx * cfg->meqn;
        ALLOC(s->reg, sizeof(double) * s->rstride * max_patches);
        if (hipMemset(s->reg, 0, sizeof(double) * s->rstride * max_patches) != hipSuccess)
            return cleanup(fail(HIPCLAW_EDEVICE, "initial device copies failed%s"));
    }
#undef ALLOC
    std::vector<int> iota(max_patches);
    for (int p = 0; p < max_patches; p++) iota[p] = p;
    if (hipMemcpy(s->all_slots, iota.data(), sizeof(int) * max_patches, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(s->q[0], 0, sizeof(double) * s->pstride * max_patches) != hipSuccess)
        return cleanup(fail(HIPCLAW_EDEVICE, "initial device copies failed%s"));
    if (s->q[1] && hipMemset(s->q[1], 0, sizeof(double) * s->pstride * max_patches) != hipSuccess)
        return cleanup(fail(HIPCLAW_EDEVICE, "initial device copies failed%s"));
    if (hipEventCreate(&s->region[0]) != hipSuccess || hipEventCreate(&s->region[1]) != hipSuccess)
        return cleanup(fail(HIPCLAW_EDEVICE, "hipEventCreate failed%s"));
    *out = s;
    return HIPCLAW_OK;
}

void hipclaw_destroy(hipclaw_state_t* s) {
    if (!s) return;
    if (s->st) hipStreamSynchronize(s->st);
    if (s->gexec) hipGraphExecDestroy(s->gexec);
    if (s->graph) hipGraphDestroy(s->graph);
    for (hipEvent_t e : s->kev) hipEventDestroy(e);
    for (hipEvent_t e : s->region) if (e) hipEventDestroy(e);
    hipFree(s->q[0]);
    hipFree(s->q[1]);
    hipFree(s->aux);
    hipFree(s->dxs);
    hipFree(s->dys);
    hipFree(s->all_slots);
    hipFree(s->batch_slots);
    for (GhostProgram& g : s->gp) {
        hipFree(g.ops);
        hipFree(g.groups);
    }
    for (SyncProgram& y : s->sp) {
        hipFree(y.ops);
        hipFree(y.start);
        hipFree(y.resets);
    }
    hipFree(s->qlast);
    hipFree(s->qti);
    hipFree(s->lslots);
    if (s->hstage) {
        if (s->st) hipStreamSynchronize(s->st);
        hipHostFree(s->hstage);
    }
    hipFree(s->qsave);
    hipFree(s->cfl);
    hipFree(s->mass_buf);
    hipFree(s->reg);
    hipFree(s->band);
    hipFree(s->dtscale);
    hipFree(s->fops);
    hipFree(s->fsops);
    hipFree(s->pull);
    hipFree(s->iwork);
    hipFree(s->frel);
    hipFree(s->cmask);
    hipFree(s->qslots);
    hipFree(s->qcorr);
    if (s->st) hipStreamDestroy(s->st);
    delete s;
}

int hipclaw_tile(const hipclaw_state_t* s) { return s ? s->tile : 0; }

int hipclaw_set_geometry(hipclaw_state_t* s, int npatch, const double* dx, const double* dy) {
    if (!s || npatch < 0 || npatch > s->max_patches || (npatch > 0 && (!dx || !dy)))
        return fail(HIPCLAW_EINVAL, "hipclaw_set_geometry: bad arguments%s");
    for (int p = 0; p < npatch; p++)
        if (!(dx[p] > 0.0) || !(dy[p] > 0.0)) return fail(HIPCLAW_EINVAL, "non-positive mesh width%s");
    if (npatch != s->npatch) drop_graph(s);
    if (npatch != s->npatch) s->lmax = s->lmin - 1;   // the levels name other slots
    s->npatch = npatch;
    if (npatch > 0) {
        HIP_TRY(hipMemcpyAsync(s->dxs, dx, sizeof(double) * npatch, hipMemcpyHostToDevice, s->st));
        HIP_TRY(hipMemcpyAsync(s->dys, dy, sizeof(double) * npatch, hipMemcpyHostToDevice, s->st));
        HIP_TRY(hipStreamSynchronize(s->st));
    }
    return HIPCLAW_OK;
}

static int check_range(hipclaw_state_t* s, const void* p, int first, int count) {
    if (!s || !p || first < 0 || count < 0 || first + count > s->max_patches)
        return fail(HIPCLAW_EINVAL, "patch range out of bounds%s");
    return HIPCLAW_OK;
}

int hipclaw_set_q(hipclaw_state_t* s, const double* q, int first, int count) {
    if (int rc = check_range(s, q, first, count)) return rc;
    if (int rc = guard_save(s)) return rc;
    const size_t bytes = sizeof(double) * s->pstride * count;
    HIP_TRY(hipMemcpyAsync(s->q[s->cur] + s->pstride * first, q, bytes, hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return HIPCLAW_OK;
}

int hipclaw_get_q(hipclaw_state_t* s, double* q, int first, int count) {
    if (int rc = check_range(s, q, first, count)) return rc;
    const size_t bytes = sizeof(double) * s->pstride * count;
    HIP_TRY(hipMemcpyAsync(q, s->q[s->cur] + s->pstride * first, bytes, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return HIPCLAW_OK;
}

int hipclaw_set_aux(hipclaw_state_t* s, const double* aux, int first, int count) {
    if (int rc = check_range(s, aux, first, count)) return rc;
    if (s->cfg.maux == 0) return HIPCLAW_OK;
    const size_t bytes = sizeof(double) * s->astride * count;
    HIP_TRY(hipMemcpyAsync(s->aux + s->astride * first, aux, bytes, hipMemcpyHostToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return HIPCLAW_OK;
}

// A slot of a ghost op: the solution, or (flag) the time-interpolated copy.
static bool op_slot_ok(const hipclaw_state_t* s, int slot) {
    const int p = slot & ~HIPCLAW_TIME_INTERP_SLOT;
    if ((slot & HIPCLAW_TIME_INTERP_SLOT) && !s->qti) return false;
    return slot >= 0 && p < s->npatch;
}

// The op types each phase of a ghost program may hold (bit mask over GHOST_*).
static const int kPhaseTypes[6] = {
    (1 << hipclaw::GHOST_AVERAGE_FACE) | (1 << hipclaw::GHOST_AVERAGE_CORNER),
    (1 << hipclaw::GHOST_COPY_FACE) | (1 << hipclaw::GHOST_COPY_CORNER),
    1 << hipclaw::GHOST_PHYSBC,
    (1 << hipclaw::GHOST_INTERP_FACE) | (1 << hipclaw::GHOST_INTERP_CORNER),
    1 << hipclaw::GHOST_PHYSBC,
    1 << hipclaw::GHOST_PHYSBC};

// Host check of one ghost op of phase ph (a bad slot or translation would
// fault the GPU).
static int check_ghost_op(const hipclaw_state_t* s, const int* op, int ph) {
    const int mx = s->cfg.mx, my = s->cfg.my, mbc = s->cfg.mbc;
    const int type = op[0], dst = op[1], src = op[2], which = op[3], ti = op[4], tj = op[5];
    // the reference's coarse/fine exchanges bound their face loops by mx
    // (fclaw2d_clawpatch46_interpolate.f, fclaw2d_clawpatch46_average.f:
    // average2coarse offsets children by mx in both directions)
    if (mx != my && type >= hipclaw::GHOST_AVERAGE_FACE && type <= hipclaw::GHOST_INTERP_CORNER)
        return fail(HIPCLAW_EINVAL, "coarse/fine ghost exchange requires mx = my%s");
    if (type < 0 || type > hipclaw::GHOST_PHYSBC || !op_slot_ok(s, dst))
        return fail(HIPCLAW_EINVAL, "ghost op with bad type or destination%s");
    if (!(kPhaseTypes[ph] & (1 << type)))
        return fail(HIPCLAW_EINVAL, "ghost op of a type its phase does not hold%s");
    if (type == hipclaw::GHOST_PHYSBC) {
        // per face: -1 (not a physical face), 1 (extrapolation) or 3 (wall).  The
        // reference stops on mthbc 0 and 2 and leaves other codes' ghost cells
        // unwritten (clawpack46_bc2_default.f:17-25, 41-43), which with two slabs
        // would leave them to whatever the spare slab held.
        for (int f = 0; f < 4; f++) {
            const int b = (int) (signed char) ((op[6] >> (8 * f)) & 0xff);
            if (b != -1 && b != 1 && b != 3)
                return fail(HIPCLAW_EINVAL,
                            "physical boundary condition code must be 1 (extrapolation) or 3 (wall)%s");
        }
        return HIPCLAW_OK;
    }
    if (!op_slot_ok(s, src) || which < 0 || which > 3)
        return fail(HIPCLAW_EINVAL, "ghost op with bad source or face%s");
    // every read index i + ti (same size) or 2*i - 1 + ti .. +1 (half size) must lie in the
    // neighbour's array; check the extreme cells the op can touch
    int lo_i, hi_i, lo_j, hi_j;  // range of the translated index
    if (type == hipclaw::GHOST_COPY_FACE || type == hipclaw::GHOST_COPY_CORNER) {
        lo_i = 1 - mbc + ti; hi_i = mx + mbc + ti; lo_j = 1 - mbc + tj; hi_j = my + mbc + tj;
        // copy reads the neighbour's interior band only
        int ci0, ci1, cj0, cj1;
        if (type == hipclaw::GHOST_COPY_FACE) {
            if (which == 0) { ci0 = 1 - mbc; ci1 = 0; cj0 = 1; cj1 = my; }
            else if (which == 1) { ci0 = mx + 1; ci1 = mx + mbc; cj0 = 1; cj1 = my; }
            else if (which == 2) { ci0 = 1; ci1 = mx; cj0 = 1 - mbc; cj1 = 0; }
            else { ci0 = 1; ci1 = mx; cj0 = my + 1; cj1 = my + mbc; }
        } else {
            ci0 = (which & 1) ? mx + 1 : 1 - mbc; ci1 = ci0 + mbc - 1;
            cj0 = (which & 2) ? my + 1 : 1 - mbc; cj1 = cj0 + mbc - 1;
        }
        lo_i = ci0 + ti; hi_i = ci1 + ti; lo_j = cj0 + tj; hi_j = cj1 + tj;
        if (lo_i < 1 || hi_i > mx || lo_j < 1 || hi_j > my)
            return fail(HIPCLAW_EINVAL, "copy op reads outside the neighbour's interior%s");
    } else if (type == hipclaw::GHOST_AVERAGE_CORNER) {
        const int ic0 = (which & 1) ? mx + 1 : 1 - mbc, jc0 = (which & 2) ? my + 1 : 1 - mbc;
        lo_i = 2 * ic0 - 1 + ti; hi_i = 2 * (ic0 + mbc - 1) + ti;
        lo_j = 2 * jc0 - 1 + tj; hi_j = 2 * (jc0 + mbc - 1) + tj;
        if (lo_i < 1 - mbc || hi_i > mx + mbc || lo_j < 1 - mbc || hi_j > my + mbc)
            return fail(HIPCLAW_EINVAL, "average-corner op reads outside the fine patch%s");
    } else if (type == hipclaw::GHOST_INTERP_CORNER) {
        const int hb = mbc / 2;
        const int ic0 = (which & 1) ? mx - hb + 1 : 1, jc0 = (which & 2) ? my - hb + 1 : 1;
        lo_i = 2 * ic0 - 1 + ti; hi_i = 2 * (ic0 + hb - 1) + ti;
        lo_j = 2 * jc0 - 1 + tj; hi_j = 2 * (jc0 + hb - 1) + tj;
        if (lo_i < 1 - mbc || hi_i > mx + mbc || lo_j < 1 - mbc || hi_j > my + mbc)
            return fail(HIPCLAW_EINVAL, "interpolate-corner op writes outside the fine patch%s");
    }
    // average/interpolate faces test every cell against the neighbour's bounds in the kernel
    return HIPCLAW_OK;
}

// The two launch layouts of ghost program k (hipclaw_set_ghost_layout 0 and 1)
// from its validated ops.
static int build_layouts(hipclaw_state_t* s, int k, const int* ops, int nops, const int* phase_start) {
    // The launches: phases 0 and 1 (average, copy: disjoint ghost cells, interior
    // reads only) share one, in the caller's order (neighbouring patches' ops
    // stay close, which keeps the cache lines they share in cache).  Phases 4
    // and 5 apply the physical BC to the fine levels, then to every level; a BC
    // touches only its own patch, so when every patch of phase 4 is in phase 5
    // with the same BC, one launch applies it twice to those patches.
    GhostProgram& g = s->gp[k];
    std::vector<GhostOp> dev;
    dev.reserve((size_t) nops);
    int layout = 0, nl[2] = {0, 0};
    auto to_op = [&](int o, int reps) {
        const int* q = ops + 8 * o;
        GhostOp op;
        op.type = q[0]; op.dst = q[1]; op.src = q[2]; op.which = q[3];
        op.ti = q[4]; op.tj = q[5]; op.bc = q[6]; op.reps = reps;
        return op;
    };
    // Within an exchange launch, a copy op and its reverse (P <- Q beside
    // Q <- P) sit next to each other, in one workgroup and so on one XCD: the
    // row-end cache lines the pair reads and writes are then shared in that
    // XCD's L2.  Other ops keep the caller's order.
    // (buckets by the lower slot of a copy, the destination otherwise; a copy
    // follows the op before it in its bucket that has the same slot pair)
    const int TI = HIPCLAW_TIME_INTERP_SLOT;
    auto add_exchange = [&](int a, int b) {
        GhostLaunch L;
        L.start = (int) dev.size();
        const int nb = 2 * s->npatch;                 // slots, then time-interpolated slots
        auto bucket = [&](int slot) { return (slot & TI) ? s->npatch + (slot & ~TI) : slot; };
        auto lo_of = [&](int o) {
            const int* q = ops + 8 * o;
            const bool copy = q[0] == hipclaw::GHOST_COPY_FACE || q[0] == hipclaw::GHOST_COPY_CORNER;
            return bucket(copy ? std::min(q[1], q[2]) : q[1]);
        };
        std::vector<int> head(nb + 1, 0), order((size_t) (b - a));
        for (int o = a; o < b; o++) head[lo_of(o) + 1]++;
        for (int i = 0; i < nb; i++) head[i + 1] += head[i];
        for (int o = a; o < b; o++) order[head[lo_of(o)]++] = o;   // stable bucket pass
        for (int i = nb; i > 0; i--) head[i] = head[i - 1];
        head[0] = 0;
        for (int i = 0; i < nb; i++) {
            // within a bucket (a few ops), bring each copy's reverse next to it
            for (int x = head[i]; x < head[i + 1]; x++) {
                const int* q = ops + 8 * order[x];
                if (q[0] != hipclaw::GHOST_COPY_FACE && q[0] != hipclaw::GHOST_COPY_CORNER) continue;
                for (int y = x + 1; y < head[i + 1]; y++) {
                    const int* r = ops + 8 * order[y];
                    if (r[0] == q[0] && r[1] == q[2] && r[2] == q[1]) {
                        const int v = order[y];
                        for (int z = y; z > x + 1; z--) order[z] = order[z - 1];
                        order[x + 1] = v;
                        x++;
                        break;
                    }
                }
            }
        }
        for (int o : order) dev.push_back(to_op(o, 1));
        L.nops = (int) dev.size() - L.start;
        if (L.nops > 0) g.launch[layout][nl[layout]++] = L;
    };
    auto add_physbc = [&](int a, int b, const std::unordered_map<int, int>* twice) {
        GhostLaunch L;
        L.start = (int) dev.size();
        L.kind = GHOST_LAUNCH_PHYSBC;
        for (int o = a; o < b; o++) dev.push_back(to_op(o, twice && twice->count(ops[8 * o + 1]) ? 2 : 1));
        L.nops = (int) dev.size() - L.start;
        if (L.nops > 0) g.launch[layout][nl[layout]++] = L;
    };
    bool merge45 = true;
    std::unordered_map<int, int> in5, in4;   // destination -> packed BC
    for (int o = phase_start[5]; o < phase_start[6] && merge45; o++)
        merge45 = in5.emplace(ops[8 * o + 1], ops[8 * o + 6]).second;
    for (int o = phase_start[4]; o < phase_start[5] && merge45; o++) {
        auto it = in5.find(ops[8 * o + 1]);
        merge45 = it != in5.end() && it->second == ops[8 * o + 6] &&
                  in4.emplace(ops[8 * o + 1], 1).second;
    }
    // Per-destination groups: the exchange ops of ops [a, b) by destination
    // patch, each group followed by that patch's physical-BC op of [pa, pb)
    // (applied `twice` where listed), in one launch.  Needs at most one BC op
    // per patch; false (nothing added) otherwise.
    std::vector<int4> groups;
    auto add_groups = [&](int a, int b, int pa, int pb, const std::unordered_map<int, int>* twice) {
        std::unordered_map<int, int> gi;   // destination -> group
        std::vector<std::vector<int>> members;
        std::vector<int> bc;
        for (int o = a; o < b; o++) {
            auto it = gi.emplace(ops[8 * o + 1], (int) members.size());
            if (it.second) { members.emplace_back(); bc.push_back(-1); }
            members[it.first->second].push_back(o);
        }
        for (int o = pa; o < pb; o++) {
            auto it = gi.emplace(ops[8 * o + 1], (int) members.size());
            if (it.second) { members.emplace_back(); bc.push_back(-1); }
            if (bc[it.first->second] >= 0) return false;
            bc[it.first->second] = o;
        }
        GhostLaunch L;
        L.kind = GHOST_LAUNCH_GROUPS;
        L.start = (int) dev.size();
        L.group0 = (int) groups.size();
        for (size_t k = 0; k < members.size(); k++) {
            int4 gr;   // op indices relative to the launch's first op
            gr.x = (int) dev.size() - L.start;
            for (int o : members[k]) dev.push_back(to_op(o, 1));
            gr.y = (int) dev.size() - L.start;
            gr.z = -1;
            gr.w = 0;
            if (bc[k] >= 0) {
                gr.z = (int) dev.size() - L.start;
                dev.push_back(to_op(bc[k], twice && twice->count(ops[8 * bc[k] + 1]) ? 2 : 1));
            }
            groups.push_back(gr);
        }
        L.nops = (int) dev.size() - L.start;
        L.ngroups = (int) groups.size() - L.group0;
        if (L.ngroups > 0) g.launch[layout][nl[layout]++] = L;
        return true;
    };
    // layout 0: one launch per phase group
    add_exchange(phase_start[0], phase_start[2]);
    add_physbc(phase_start[2], phase_start[3], nullptr);
    add_exchange(phase_start[3], phase_start[4]);
    if (merge45) {
        add_physbc(phase_start[5], phase_start[6], &in4);
    } else {
        add_physbc(phase_start[4], phase_start[5], nullptr);
        add_physbc(phase_start[5], phase_start[6], nullptr);
    }
    // layout 1, phases 0-2: averages and copies into a patch, then its BC
    // (coarse levels)
    layout = 1;
    if (!add_groups(phase_start[0], phase_start[2], phase_start[2], phase_start[3], nullptr)) {
        add_exchange(phase_start[0], phase_start[2]);
        add_physbc(phase_start[2], phase_start[3], nullptr);
    }
    // phases 3-5: interpolations into a patch, then its BC (fine levels twice)
    if (!(merge45 && add_groups(phase_start[3], phase_start[4], phase_start[5], phase_start[6], &in4))) {
        add_exchange(phase_start[3], phase_start[4]);
        if (merge45) {
            add_physbc(phase_start[5], phase_start[6], &in4);
        } else {
            add_physbc(phase_start[4], phase_start[5], nullptr);
            add_physbc(phase_start[5], phase_start[6], nullptr);
        }
    }
    if (int rc = ensure_dev(g.ops, g.cap, (int) dev.size(), "ghost ops")) return rc;
    if (!dev.empty()) {
        HIP_TRY(hipMemcpyAsync(g.ops, dev.data(), sizeof(GhostOp) * dev.size(), hipMemcpyHostToDevice, s->st));
        HIP_TRY(hipStreamSynchronize(s->st));
    }
    if (int rc = ensure_dev(g.groups, g.groups_cap, (int) groups.size(), "ghost groups")) return rc;
    if (!groups.empty()) {
        HIP_TRY(hipMemcpyAsync(g.groups, groups.data(), sizeof(int4) * groups.size(), hipMemcpyHostToDevice, s->st));
        HIP_TRY(hipStreamSynchronize(s->st));
    }
    g.nops = (int) dev.size();
    g.nlaunch[0] = nl[0];
    g.nlaunch[1] = nl[1];
    g.layouts_built = true;
    return HIPCLAW_OK;
}

// Layouts 0 and 1 of program 0, built on first use (never inside a graph
// capture: the callers that may capture call this first).
static int ensure_layouts(hipclaw_state_t* s) {
    GhostProgram& g = s->gp[0];
    if (g.layouts_built || !s->gp0_set) return HIPCLAW_OK;
    return build_layouts(s, 0, s->gp0_ops.data(), (int) s->gp0_ops.size() / 8, s->gp0_phase);
}

int hipclaw_set_ghost_program_k(hipclaw_state_t* s, int k, const int* ops, int nops,
                                const int* phase_start) {
    if (!s || k < 0 || k >= HIPCLAW_MAX_PROGRAMS || nops < 0 || (nops > 0 && !ops) || !phase_start)
        return fail(HIPCLAW_EINVAL, "hipclaw_set_ghost_program: bad arguments%s");
    if (phase_start[0] != 0 || phase_start[6] != nops)
        return fail(HIPCLAW_EINVAL, "phase_start must run from 0 to nops%s");
    for (int k = 0; k < 6; k++)
        if (phase_start[k + 1] < phase_start[k]) return fail(HIPCLAW_EINVAL, "phase_start not monotone%s");
    // every op is checked on the host before anything of the state changes (a
    // bad slot or translation would fault the GPU); program 0's face table is
    // built in the same pass
    FusedBuild fb;
    if (k == 0) {
        if (int rc = fused_check_and_faces(s, ops, nops, phase_start, fb)) return rc;
    } else {
        int ph = 0;
        for (int o = 0; o < nops; o++) {
            while (o >= phase_start[ph + 1]) ph++;
            if (int rc = check_ghost_op(s, ops + 8 * o, ph)) return rc;
        }
    }
    // a captured graph holds the tables about to be replaced
    drop_graph(s);
    GhostProgram& g = s->gp[k];
    g.layouts_built = false;
    g.set = true;
    g.nops = 0;
    g.nlaunch[0] = g.nlaunch[1] = 0;
    if (k == 0) {
        s->gp0_ops.assign(ops, ops + 8 * (size_t) nops);
        for (int p = 0; p < 7; p++) s->gp0_phase[p] = phase_start[p];
        s->gp0_set = true;
        if (int rc = fused_finish(s, ops, nops, fb)) return rc;
        // the fused step needs no launch layout of its own; the two-launch
        // layouts are built when one is asked for (ensure_layouts)
        if (s->ghost_layout == 2 && s->fused_ok) return HIPCLAW_OK;
    }
    return build_layouts(s, k, ops, nops, phase_start);
}

int hipclaw_set_ghost_program(hipclaw_state_t* s, const int* ops, int nops, const int* phase_start) {
    return hipclaw_set_ghost_program_k(s, 0, ops, nops, phase_start);
}

static int upload(hipclaw_state_t* s, int*& dev, int& cap, const int* host, int n, const char* what) {
    if (int rc = ensure_dev(dev, cap, n, what)) return rc;
    if (n > 0) HIP_TRY(hipMemcpyAsync(dev, host, sizeof(int) * (size_t) n, hipMemcpyHostToDevice, s->st));
    return HIPCLAW_OK;
}

int hipclaw_set_sync_program_k(hipclaw_state_t* s, int k, const int* ops, int nops,
                               const int* resets, int nresets) {
    if (!s || k < 0 || k >= HIPCLAW_MAX_PROGRAMS || nops < 0 || (nops > 0 && !ops) ||
        nresets < 0 || (nresets > 0 && !resets))
        return fail(HIPCLAW_EINVAL, "hipclaw_set_sync_program: bad arguments%s");
    if (!s->reg && (nops > 0 || nresets > 0))
        return fail(HIPCLAW_EINVAL, "the state was created without time_sync%s");
    // validate on the host and group by destination (groups must be contiguous)
    std::vector<int> start;
    std::vector<char> seen(s->npatch, 0);
    for (int o = 0; o < nops; o++) {
        const int* op = ops + 8 * o;
        const int type = op[0], dst = op[1], src = op[2], face = op[3];
        if ((type != hipclaw::SYNC_SAMESIZE && type != hipclaw::SYNC_F2C) || dst < 0 ||
            dst >= s->npatch || src < 0 || src >= s->npatch || face < 0 || face > 3)
            return fail(HIPCLAW_EINVAL, "sync op with bad type, slot or face%s");
        if (o == 0 || op[1] != ops[8 * (o - 1) + 1]) {
            if (seen[dst]) return fail(HIPCLAW_EINVAL, "sync ops of one destination must be contiguous%s");
            seen[dst] = 1;
            start.push_back(o);
        }
    }
    start.push_back(nops);
    for (int r = 0; r < nresets; r++)
        if (resets[2 * r] < 0 || resets[2 * r] >= s->npatch || resets[2 * r + 1] < 0 || resets[2 * r + 1] > 3)
            return fail(HIPCLAW_EINVAL, "register reset with bad slot or face%s");
    drop_graph(s);
    SyncProgram& y = s->sp[k];
    const int ngroups = (int) start.size() - 1;
    if (int rc = upload(s, y.ops, y.ops_cap, ops, 8 * nops, "sync ops")) return rc;
    if (int rc = upload(s, y.start, y.groups_cap, start.data(), ngroups + 1, "sync groups")) return rc;
    if (int rc = upload(s, y.resets, y.resets_cap, resets, 2 * nresets, "register resets")) return rc;
    HIP_TRY(hipStreamSynchronize(s->st));
    y.ngroups = ngroups;
    y.nresets = nresets;
    y.maxgroup = 0;
    y.set = true;
    for (int g = 0; g < ngroups; g++) y.maxgroup = std::max(y.maxgroup, start[g + 1] - start[g]);
    if (k == 0) {
        s->sp0_ops.assign(ops, ops + 8 * (size_t) nops);
        return build_fused_sync(s);
    }
    return HIPCLAW_OK;
}

}  // extern "C"

// Pass one over program 0: every op checked (nothing of the state changes
// before all have passed), the face table filled in the staging buffer, and
// the ops of each part of the flat list counted.
static int fused_check_and_faces(hipclaw_state_t* s, const int* ops, int nops, const int* phase_start,
                                 FusedBuild& b) {
    using hipclaw::FaceRel;
    const int np = s->npatch;
    const int TI = HIPCLAW_TIME_INTERP_SLOT;
    b.fusable = np > 0;
    FaceRel* frel = nullptr;
    if (b.fusable) {
        const size_t frel_bytes = sizeof(FaceRel) * 4 * (size_t) np;
        b.off_fops = (frel_bytes + 255) & ~(size_t) 255;
        b.off_pull = b.off_fops + ((sizeof(GhostOp) * (size_t) nops + 255) & ~(size_t) 255);
        if (int rc = stage_reserve(s, b.off_pull + sizeof(int) * 8 * (size_t) np, &b.stage)) return rc;
        frel = reinterpret_cast<FaceRel*>(b.stage);
        for (size_t k = 0; k < 4 * (size_t) np; k++) frel[k] = FaceRel{hipclaw::FREL_NONE, -1, 0, 0, -1, 0, 0, 0};
        b.bcop.assign((size_t) np, -1);
    }
    int ph = 0;
    for (int o = 0; o < nops; o++) {
        const int* op = ops + 8 * o;
        while (o >= phase_start[ph + 1]) ph++;
        if (int rc = check_ghost_op(s, op, ph)) return rc;
        if (!b.fusable) continue;
        const int type = op[0], dst = op[1], src = op[2], which = op[3];
        // a program with time-interpolated slots stays out of the fused step
        if ((dst & TI) || (type != hipclaw::GHOST_PHYSBC && (src & TI))) {
            b.fusable = false;
            continue;
        }
        if (type == hipclaw::GHOST_COPY_FACE) {
            frel[4 * dst + which] = FaceRel{hipclaw::FREL_SAME, src, op[4], op[5], -1, 0, 0, 0};
            b.ncopy++;
        } else if (type == hipclaw::GHOST_AVERAGE_FACE) {
            FaceRel& r = frel[4 * dst + which];
            if (r.rel != hipclaw::FREL_HALF) r = FaceRel{hipclaw::FREL_HALF, src, op[4], op[5], -1, 0, 0, 0};
            else { r.nb1 = src; r.ti1 = op[4]; r.tj1 = op[5]; }
            b.nlite++;
        } else if (type == hipclaw::GHOST_INTERP_FACE) {
            frel[4 * dst + (which ^ 1)] = FaceRel{hipclaw::FREL_DOUBLE, src, op[4], op[5], -1, 0, 0, which};
            b.nlite++;
        } else if (type == hipclaw::GHOST_PHYSBC) {
            // one op per patch in the flat list: every op of a patch must hold the same codes
            const int prev = b.bcop[dst];
            if (prev >= 0 && ops[8 * prev + 6] != op[6]) {
                b.fusable = false;
                continue;
            }
            if (prev < 0) b.nlite++;
            b.bcop[dst] = o;
            for (int f = 0; f < 4; f++) {
                const int c = (int) (signed char) ((op[6] >> (8 * f)) & 0xff);
                if (c == 1 || c == 3) frel[4 * dst + f] = FaceRel{hipclaw::FREL_PHYS, -1, 0, 0, -1, 0, 0, c};
            }
        } else if (type == hipclaw::GHOST_COPY_CORNER) {
            b.ncopyc++;
        } else {
            b.nlite++;   // average and interpolate corners
            b.ncorner++;
        }
    }
    return HIPCLAW_OK;
}

// The fused step of program 0 (ghost layout 2).  Every ghost op becomes one
// wave of ghost_flat (a corner op a quarter of one) that reads interior cells
// only: copies and averages do
// already; interpolation and the physical BC recompute the ghost cells they
// read from the face table (same-size neighbour, fine children, physical BC,
// or for a physical BC beside a face filled from a coarser patch, that
// interpolation).  The program is fused only if every such cell has one of
// those sources (which the 2:1 balance of the forest guarantees) and it has no
// time-interpolated slots; otherwise fused_ok stays false and the advance
// keeps the two-launch layouts.  With two slabs the update also stages the
// same-size ghost strips and corners straight from the neighbours (pull), and
// the copy ops run only in the last step of an advance, so that the slab's
// ghost cells are complete for readers.
static int fused_finish(hipclaw_state_t* s, const int* ops, int nops, FusedBuild& b) {
    using hipclaw::FaceRel;
    s->fused_ok = false;
    s->fsync = false;
    s->pull_ok = false;
    if (!b.fusable) return HIPCLAW_OK;
    const int np = s->npatch, mx = s->cfg.mx, my = s->cfg.my, mbc = s->cfg.mbc;
    const int nxp = mx + 2 * mbc;
    const FaceRel* frel = reinterpret_cast<const FaceRel*>(b.stage);
    // the face table travels while the host builds the other two
    if (int rc = stage_copy(s, s->frel, s->frel_cap, frel, 4 * np, "face table")) return rc;
    // cells read from a coarse patch's face strips (interpolation) need a source
    // that is not itself an interpolation; a physical BC may read any face strip
    auto simple = [&](int c, int f) {
        const FaceRel& r = frel[4 * c + f];
        return r.rel == hipclaw::FREL_SAME || r.rel == hipclaw::FREL_PHYS ||
               (r.rel == hipclaw::FREL_HALF && r.nb1 >= 0);
    };
    auto to_op = [&](const int* q) {
        GhostOp op;
        op.type = q[0]; op.dst = q[1]; op.src = q[2]; op.which = q[3];
        op.ti = q[4]; op.tj = q[5]; op.bc = q[6]; op.reps = 1;
        return op;
    };
    // flat op list: the corners of every step, its other ops, the face copies,
    // the corner copies; the same-size strips and corners the update stages
    // from the neighbours (pull offsets in doubles must fit an int)
    GhostOp* fops = reinterpret_cast<GhostOp*>(b.stage + b.off_fops);
    int* pull = reinterpret_cast<int*>(b.stage + b.off_pull);
    const bool pull_ok = (double) s->pstride * s->max_patches < 2.0e9;
    if (pull_ok) std::fill(pull, pull + 8 * (size_t) np, 0);
    int nc = 0, nl = b.ncorner, nf = b.nlite, nfc = b.nlite + b.ncopy;
    for (int o = 0; o < nops; o++) {
        const int* op = ops + 8 * o;
        const int type = op[0];
        if (type == hipclaw::GHOST_COPY_FACE || type == hipclaw::GHOST_COPY_CORNER) {
            fops[type == hipclaw::GHOST_COPY_FACE ? nf++ : nfc++] = to_op(op);
            if (pull_ok) {
                const long long d = ((long long) op[2] - op[1]) * s->pstride + op[4] + (long long) op[5] * nxp;
                pull[8 * (size_t) op[1] + (type == hipclaw::GHOST_COPY_FACE ? 0 : 4) + op[3]] = (int) d;
            }
            continue;
        }
        const int c = op[2], which = op[3], ti = op[4], tj = op[5];
        if (type == hipclaw::GHOST_INTERP_FACE) {
            if (!(frel[4 * c + which].rel == hipclaw::FREL_HALF && frel[4 * c + which].nb1 >= 0)) return HIPCLAW_OK;
            // whether the op writes the fine cells at either end of the face (the
            // kernel bounds its loop by mx, mx == my, and skips fine cells
            // outside the patch)
            auto writes = [&](int l) {
                const int ic = which < 2 ? (which == 0 ? 1 : mx) : l, jc = which < 2 ? l : (which == 2 ? 1 : my);
                const int fi = 2 * ic - 1 + ti, fj = 2 * jc - 1 + tj;
                return !(fi < 1 - mbc || fi + 1 > mx + mbc || fj < 1 - mbc || fj + 1 > my + mbc);
            };
            const int flo = which < 2 ? 2 : 0, fhi = which < 2 ? 3 : 1;
            if ((writes(1) && !simple(c, flo)) || (writes(mx) && !simple(c, fhi))) return HIPCLAW_OK;
        } else if (type == hipclaw::GHOST_INTERP_CORNER) {
            if (!simple(c, (which & 1) ? 1 : 0) || !simple(c, (which & 2) ? 3 : 2)) return HIPCLAW_OK;
        } else if (type == hipclaw::GHOST_PHYSBC) {
            for (int f = 0; f < 4; f++)
                if (frel[4 * op[1] + f].rel == hipclaw::FREL_NONE) return HIPCLAW_OK;
            if (b.bcop[op[1]] != o) continue;   // once per patch
        }
        const bool corner = type == hipclaw::GHOST_AVERAGE_CORNER || type == hipclaw::GHOST_INTERP_CORNER;
        fops[corner ? nc++ : nl++] = to_op(op);
    }
    if (nc != b.ncorner || nl != b.nlite || nf != b.nlite + b.ncopy || nfc != nf + b.ncopyc)
        return fail(HIPCLAW_EINVAL, "internal: fused ghost op count mismatch%s");
    if (int rc = stage_copy(s, s->fops, s->fops_cap, fops, nfc, "fused ghost ops")) return rc;
    if (pull_ok)
        if (int rc = stage_copy(s, s->pull, s->pull_cap, pull, 8 * np, "pull offsets")) return rc;
    // the copies run on while the host goes on (the next user of the staging
    // buffer waits for them; launches on the stream come after them)
    s->fc0 = b.ncorner;
    s->fbig = b.nlite - b.ncorner;
    s->fcopy = b.ncopy;
    s->fc1 = b.ncopyc;
    s->pull_ok = pull_ok;
    s->fused_ok = true;
    return build_fused_sync(s);
}

// The time-sync part of the fused step: the faces each patch corrects (from
// sync program 0, contiguous per destination) and room for the increments.
static int build_fused_sync(hipclaw_state_t* s) {
    s->fsync = false;
    if (!s->fused_ok) return HIPCLAW_OK;
    const int np = s->npatch, mx = s->cfg.mx;
    const bool sync = s->reg && !s->reg_accumulate && !s->sp0_ops.empty();
    const int nso = (int) s->sp0_ops.size() / 8;
    const int* ops = s->sp0_ops.data();
    if (sync)
        for (int o = 0; o < nso; o++)
            if (ops[8 * o + 1] < 0 || ops[8 * o + 1] >= np || ops[8 * o + 2] < 0 || ops[8 * o + 2] >= np)
                return HIPCLAW_OK;
    s->nfsync = 0;
    s->nqslots = 0;
    s->fpaired = false;
    if (!sync) return HIPCLAW_OK;
    // the faces each patch corrects and, in the FP32 flux form, the paired
    // sync list, built in the pinned staging buffer and copied out together
    const bool pair = s->cfg.precision == HIPCLAW_FP32_BULK && s->conservative;
    const size_t off_fs = ((sizeof(int) * (size_t) np) + 255) & ~(size_t) 255;
    const size_t off_qs = off_fs + (((pair ? sizeof(int) * 8 * (size_t) nso : 0) + 255) & ~(size_t) 255);
    char* stage = nullptr;
    if (int rc = stage_reserve(s, off_qs + sizeof(int) * (size_t) np, &stage)) return rc;
    int* cm = reinterpret_cast<int*>(stage);
    std::fill(cm, cm + np, 0);
    for (int o = 0; o < nso; o++) {
        const int* op = ops + 8 * o;
        cm[op[1]] |= (1 << op[3]) | (op[0] == hipclaw::SYNC_F2C ? 1 << (4 + op[3]) : 0);
    }
    if (pair) {
        // one wave per same-size face pair, the op with the smaller
        // (dst, face) key writing both increments
        int* fs = reinterpret_cast<int*>(stage + off_fs);
        int nf = 0;
        std::vector<int> same(4 * (size_t) np, -1);   // same-size op of (dst, face), -1 if none
        for (int o = 0; o < nso; o++)
            if (ops[8 * o] == hipclaw::SYNC_SAMESIZE) same[4 * (size_t) ops[8 * o + 1] + ops[8 * o + 3]] = o;
        for (int o = 0; o < nso; o++) {
            const int* op = ops + 8 * o;
            int type = op[0];
            if (type == hipclaw::SYNC_SAMESIZE) {
                const long long key = 4LL * op[1] + op[3], pkey = 4LL * op[2] + (op[3] ^ 1);
                const int q = same[(size_t) pkey];
                if (q >= 0 && ops[8 * q + 2] == op[1]) {
                    if (pkey < key) continue;          // the partner writes this one
                    type = hipclaw::SYNC_SAMESIZE_PAIR;
                }
            }
            int* d = fs + 8 * (size_t) nf++;
            std::memcpy(d, op, 8 * sizeof(int));
            d[0] = type;
        }
        if (int rc = stage_copy(s, s->fsops, s->fsops_cap, fs, 8 * nf, "fused sync ops")) return rc;
        s->nfsync = nf;
        s->fpaired = true;
    }
    if (int rc = stage_copy(s, s->cmask, s->cmask_cap, cm, np, "sync face masks")) return rc;
    {
        int* qs = reinterpret_cast<int*>(stage + off_qs);
        int nq = 0;
        for (int p = 0; p < np; p++)
            if (cm[p] & 0xf) qs[nq++] = p;
        if (int rc = stage_copy(s, s->qslots, s->qslots_cap, qs, nq, "corrected patches")) return rc;
        s->nqslots = nq;
    }
    {
        const long long cstride = 4LL * s->cfg.meqn * mx;
        if (!s->qcorr || s->cstride != cstride || s->qcorr_cap < s->max_patches) {
            hipFree(s->qcorr);
            s->qcorr = nullptr;
            s->qcorr_cap = 0;
            if (hipMalloc((void**) &s->qcorr, sizeof(double) * (size_t) cstride * (size_t) s->max_patches) != hipSuccess)
                return fail(HIPCLAW_ENOMEM, "hipMalloc of the time-sync increments failed%s");
            s->cstride = cstride;
            s->qcorr_cap = s->max_patches;
        }
    }
    s->fsync = sync;
    return HIPCLAW_OK;
}

extern "C" {

int hipclaw_set_sync_program(hipclaw_state_t* s, const int* ops, int nops) {
    return hipclaw_set_sync_program_k(s, 0, ops, nops, nullptr, 0);
}

static int time_sync_async(hipclaw_state_t* s, int k) {
    const SyncProgram& y = s->sp[k];
    if (!s->reg || (y.ngroups == 0 && y.nresets == 0)) return HIPCLAW_OK;
    hipclaw::SyncArgs a;
    a.ops = y.ops;
    a.start = y.start;
    a.ngroups = y.ngroups;
    a.maxgroup = y.maxgroup;
    a.resets = y.resets;
    a.nresets = y.nresets;
    a.q = s->q[s->cur];
    a.reg = s->reg;
    a.dxs = s->dxs;
    a.dys = s->dys;
    a.pstride = s->pstride;
    a.rstride = s->rstride;
    a.mx = s->cfg.mx; a.mbc = s->cfg.mbc; a.meqn = s->cfg.meqn;
    a.fluxreg = s->cfg.precision == HIPCLAW_FP32_BULK && s->conservative;
    if (int e = hipclaw::launch_time_sync(a, s->st)) return hipfail((hipError_t) e, "time_sync launch");
    return HIPCLAW_OK;
}

int hipclaw_time_sync_k(hipclaw_state_t* s, int k) {
    if (!s || k < 0 || k >= HIPCLAW_MAX_PROGRAMS) return fail(HIPCLAW_EINVAL, "hipclaw_time_sync: bad arguments%s");
    if (int rc = guard_save(s)) return rc;
    if (int rc = time_sync_async(s, k)) return rc;
    HIP_TRY(hipStreamSynchronize(s->st));
    return HIPCLAW_OK;
}

int hipclaw_time_sync(hipclaw_state_t* s) { return hipclaw_time_sync_k(s, 0); }

// The fused launch of program 0 on the current slab: the ghost fill (`full`:
// with the same-size copies, else without the ghost cells the update pulls),
// and with `sync` the time-sync increments into qcorr.
static int ghost_flat_async(hipclaw_state_t* s, bool full, bool sync) {
    hipclaw::FlatArgs f;
    f.g.ops = s->fops;
    f.g.nops = 0;
    f.g.q = s->q[s->cur];
    f.g.qti = nullptr;
    f.g.pstride = s->pstride;
    f.g.mx = s->cfg.mx; f.g.my = s->cfg.my; f.g.mbc = s->cfg.mbc; f.g.meqn = s->cfg.meqn;
    f.ncorner0 = s->fc0;
    f.nbig = full ? s->fbig + s->fcopy : s->fbig;
    f.ncorner1 = full ? s->fc1 : 0;
    f.frel = s->frel;
    f.sops = !sync ? nullptr : s->fpaired ? s->fsops : s->sp[0].ops;
    f.nsync = !sync ? 0 : s->fpaired ? s->nfsync : (int) s->sp0_ops.size() / 8;
    f.reg = s->reg;
    f.dxs = s->dxs;
    f.dys = s->dys;
    f.rstride = s->rstride;
    f.qinc = s->qcorr;
    f.cstride = s->cstride;
    f.fluxreg = s->cfg.precision == HIPCLAW_FP32_BULK && s->conservative;
    if (int e = hipclaw::launch_ghost_flat(f, s->st)) return hipfail((hipError_t) e, "ghost_flat launch");
    return HIPCLAW_OK;
}

static int ghost_fill_async(hipclaw_state_t* s, int k) {
    const GhostProgram& p = s->gp[k];
    if (k == 0 && s->ghost_layout == 2 && s->fused_ok) return ghost_flat_async(s, true, false);
    if (k == 0 && !p.layouts_built) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIP_TRY(hipStreamIsCapturing(s->st, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(HIPCLAW_EINVAL, "ghost-fill layouts not built before the capture%s");
        if (int rc = ensure_layouts(s)) return rc;
    }
    hipclaw::GhostArgs g;
    g.q = s->q[s->cur];
    g.qti = s->qti;
    g.pstride = s->pstride;
    g.mx = s->cfg.mx; g.my = s->cfg.my; g.mbc = s->cfg.mbc; g.meqn = s->cfg.meqn;
    const int lay = s->ghost_layout == 2 ? (s->cfg.mx >= 32 ? 1 : 0) : s->ghost_layout;
    for (int l = 0; l < p.nlaunch[lay]; l++) {
        const GhostLaunch& L = p.launch[lay][l];
        g.ops = p.ops + L.start;
        g.nops = L.nops;
        const int e = L.kind == GHOST_LAUNCH_PHYSBC ? hipclaw::launch_ghost_physbc(g, s->st)
                      : L.kind == GHOST_LAUNCH_GROUPS
                          ? hipclaw::launch_ghost_groups(g, p.groups + L.group0, L.ngroups, s->st)
                          : hipclaw::launch_ghost_exchange(g, s->st);
        if (e) return hipfail((hipError_t) e, "ghost fill launch");
    }
    return HIPCLAW_OK;
}

int hipclaw_ghost_fill_k(hipclaw_state_t* s, int k) {
    if (!s || k < 0 || k >= HIPCLAW_MAX_PROGRAMS) return fail(HIPCLAW_EINVAL, "hipclaw_ghost_fill: bad arguments%s");
    if (int rc = guard_save(s)) return rc;
    if (int rc = ghost_fill_async(s, k)) return rc;
    HIP_TRY(hipStreamSynchronize(s->st));
    return HIPCLAW_OK;
}

int hipclaw_ghost_fill(hipclaw_state_t* s) { return hipclaw_ghost_fill_k(s, 0); }

// ---------------------------------------------------------------- ghost-patch messages
// refratio 2: the band keeps mint = 2 mbc interior layers
// (clawpatch_ghost_pack_elems, src/patches/clawpatch/fclaw2d_clawpatch.cpp:1020-1060)
static void message_shape(const hipclaw_state_t* s, long long* nband, long long* nreg) {
    const int mx = s->cfg.mx, my = s->cfg.my, mbc = s->cfg.mbc, mint = 2 * mbc;
    *nband = (long long) (mx + 2 * mbc) * (my + 2 * mbc) - (long long) (mx - 2 * mint) * (my - 2 * mint);
    *nreg = s->reg ? 2LL * (4 * s->cfg.meqn + 2) * (mx + my) : 0;
}

int hipclaw_ghost_pack_elems(const hipclaw_state_t* s, long long* elems) {
    if (!s || !elems) return fail(HIPCLAW_EINVAL, "hipclaw_ghost_pack_elems: bad arguments%s");
    long long nband, nreg;
    message_shape(s, &nband, &nreg);
    *elems = nband * s->cfg.meqn + nreg;
    return HIPCLAW_OK;
}

static int ghost_message(hipclaw_state_t* s, const int* slots, int n, int time_interp,
                         double* buf, int unpack) {
    const char* what = unpack ? "hipclaw_ghost_unpack" : "hipclaw_ghost_pack";
    if (!s || n < 0 || n > s->max_patches || (n > 0 && (!slots || !buf)))
        return fail(HIPCLAW_EINVAL, "ghost message: bad arguments%s");
    if (time_interp && !s->qti) return fail(HIPCLAW_EINVAL, "time_interp needs a state created with subcycle%s");
    for (int k = 0; k < n; k++)
        if (slots[k] < 0 || slots[k] >= s->npatch) return fail(HIPCLAW_EINVAL, "slot out of range%s");
    if (n == 0) return HIPCLAW_OK;
    long long nband, nreg;
    message_shape(s, &nband, &nreg);
    const long long elems = nband * s->cfg.meqn + nreg;
    // the buffer must be device memory holding n messages
    hipPointerAttribute_t attr;
    if (hipPointerGetAttributes(&attr, buf) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void) hipGetLastError();
        return fail(HIPCLAW_EINVAL, "the message buffer is not device memory%s");
    }
    void* base = nullptr;
    size_t bytes = 0;
    HIP_TRY(hipMemGetAddressRange((hipDeviceptr_t*) &base, &bytes, buf));
    if ((const char*) buf + sizeof(double) * (size_t) (elems * n) > (const char*) base + bytes)
        return fail(HIPCLAW_EINVAL, "the message buffer is too small%s");
    if (!s->band) {
        // offsets of one field's band in the order of the reference's pack loops
        const int mx = s->cfg.mx, my = s->cfg.my, mbc = s->cfg.mbc, ng = mbc, mint = 2 * mbc;
        const int nxp = mx + 2 * mbc;
        std::vector<int> off;
        auto put = [&](int i, int j) { off.push_back((j + mbc - 1) * nxp + i + mbc - 1); };
        for (int j = 1 - ng; j <= my - mint; j++) for (int i = 1 - ng; i <= mint; i++) put(i, j);
        for (int j = 1 - ng; j <= mint; j++) for (int i = mint + 1; i <= mx + ng; i++) put(i, j);
        for (int j = mint + 1; j <= my + ng; j++) for (int i = mx - mint + 1; i <= mx + ng; i++) put(i, j);
        for (int j = my - mint + 1; j <= my + ng; j++) for (int i = 1 - ng; i <= mx - mint; i++) put(i, j);
        if ((long long) off.size() != nband) return fail(HIPCLAW_EDEVICE, "ghost band size mismatch%s");
        HIP_TRY(hipMalloc((void**) &s->band, sizeof(int) * off.size()));
        HIP_TRY(hipMemcpy(s->band, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice));
        s->nband = (int) off.size();
    }
    HIP_TRY(hipMemcpyAsync(s->batch_slots, slots, sizeof(int) * n, hipMemcpyHostToDevice, s->st));
    double* q = time_interp ? s->qti : s->q[s->cur];
    int e = hipclaw::launch_ghost_message(q, s->reg, s->batch_slots, n, s->band, s->nband,
                                          s->pstride, s->rstride, s->cfg.mx, s->cfg.my,
                                          s->cfg.mbc, s->cfg.meqn, s->dxs, s->dys, buf, elems,
                                          unpack, s->st);
    if (e) return hipfail((hipError_t) e, what);
    HIP_TRY(hipStreamSynchronize(s->st));
    return HIPCLAW_OK;
}

int hipclaw_ghost_pack(hipclaw_state_t* s, const int* slots, int n, int time_interp, double* buf) {
    return ghost_message(s, slots, n, time_interp, buf, 0);
}

int hipclaw_ghost_unpack(hipclaw_state_t* s, const int* slots, int n, int time_interp,
                         const double* buf) {
    if (s && !time_interp)
        if (int rc = guard_save(s)) return rc;
    return ghost_message(s, slots, n, time_interp, const_cast<double*>(buf), 1);
}

int hipclaw_save_step(hipclaw_state_t* s) {
    if (!s) return fail(HIPCLAW_EINVAL, "hipclaw_save_step: bad arguments%s");
    if (!s->inplace) {   // fclaw2d_clawpatch_save_step without the copy
        s->save_lazy = true;
        s->save_slab = s->cur;
        s->saved = true;
        return HIPCLAW_OK;
    }
    s->save_lazy = false;
    if (int rc = copy_to_save(s, s->cur)) return rc;
    s->saved = true;
    return HIPCLAW_OK;
}

int hipclaw_restore_step(hipclaw_state_t* s) {
    if (!s) return fail(HIPCLAW_EINVAL, "hipclaw_restore_step: bad arguments%s");
    if (!s->saved) return fail(HIPCLAW_EINVAL, "no saved state for the current patches%s");
    if (s->save_lazy) {
        // the saved slab was not written since: it becomes the current one again
        s->cur = s->save_slab;
        HIP_TRY(hipStreamSynchronize(s->st));
        return HIPCLAW_OK;
    }
    const size_t bytes = sizeof(double) * (size_t) s->pstride * (size_t) s->npatch;
    if (bytes > 0) HIP_TRY(hipMemcpyAsync(s->q[s->cur], s->qsave, bytes, hipMemcpyDeviceToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return HIPCLAW_OK;
}

static int advance_launches(hipclaw_state_t* s, double dt, int nsteps, bool timed);
static int subcycle_launches(hipclaw_state_t* s, double dt, int ncoarse);
static int subcycle_ready(hipclaw_state_t* s);

// Tile sizes a state can switch to: supported by its solver and precision,
// no larger than the patch, and (FP64 with time sync) leaving LDS for the
// registers.
static bool tile_allowed(const hipclaw_state_t* s, int t) {
    return solver_tile_ok(s->cfg.solver, s->cfg.precision, t) && t <= s->cfg.mx &&
           t <= s->cfg.my && !(s->cfg.time_sync && s->cfg.precision == HIPCLAW_FP64 && t == 32);
}

static int set_tile(hipclaw_state_t* s, int t) {
    const int ntx = tiles_along(s->cfg.mx, t), nty = tiles_along(s->cfg.my, t);
    const bool inplace = ntx == 1 && nty == 1;
    if (!inplace && !s->q[1]) {   // a ping-pong tile needs the second slab
        if (hipMalloc((void**) &s->q[1], sizeof(double) * s->pstride * (size_t) s->max_patches) != hipSuccess)
            return fail(HIPCLAW_ENOMEM, "hipMalloc of the second slab failed%s");
    }
    s->tile = t;
    s->ntx = ntx;
    s->nty = nty;
    s->inplace = inplace;
    // a captured graph holds launches of the old tile
    if (s->gexec) { hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { hipGraphDestroy(s->graph); s->graph = nullptr; }
    s->graph_steps = 0;
    return HIPCLAW_OK;
}

int hipclaw_set_ghost_layout(hipclaw_state_t* s, int layout) {
    if (!s || layout < 0 || layout > 2) return fail(HIPCLAW_EINVAL, "hipclaw_set_ghost_layout: bad arguments%s");
    if (layout != 2 || !s->fused_ok)
        if (int rc = ensure_layouts(s)) return rc;
    s->ghost_layout = layout;
    // a captured graph holds the launches of the old layout
    if (s->gexec) { hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { hipGraphDestroy(s->graph); s->graph = nullptr; }
    s->graph_steps = 0;
    return HIPCLAW_OK;
}

int hipclaw_ghost_layout(const hipclaw_state_t* s) {
    if (!s) return -1;
    if (s->ghost_layout == 2 && !s->fused_ok) return s->cfg.mx >= 32 ? 1 : 0;
    return s->ghost_layout;
}

static int effective_waves(const hipclaw_state_t* s) {
    if (s->waves) return s->waves;
    return solver_default_waves(s->cfg.solver, s->cfg.precision, s->tile);
}

int hipclaw_set_waves(hipclaw_state_t* s, int waves) {
    if (!s || (waves != 0 && waves != 2 && waves != 3 && waves != 4))
        return fail(HIPCLAW_EINVAL, "hipclaw_set_waves: waves must be 0, 2, 3 or 4%s");
    s->waves = waves;
    // a captured graph holds the launches of the old kernel
    if (s->gexec) { hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { hipGraphDestroy(s->graph); s->graph = nullptr; }
    s->graph_steps = 0;
    return HIPCLAW_OK;
}

int hipclaw_waves(const hipclaw_state_t* s) { return s ? effective_waves(s) : -1; }

int hipclaw_autotune(hipclaw_state_t* s, double dt, int nsteps, int* tiles, int* waves, int* layouts,
                     double* ms, int* ntrials) {
    if (!s || nsteps <= 0 || !(dt > 0.0) || !tiles || !waves || !layouts || !ms || !ntrials)
        return fail(HIPCLAW_EINVAL, "hipclaw_autotune: bad arguments%s");
    if (int rc = guard_save(s)) return rc;
    // a subcycling state with its levels set is tuned on its subcycled coarse
    // steps (each ends with the full synchronisation, which clears the
    // registers the substeps accumulate)
    const bool sub = s->qlast && s->lmax >= s->lmin;
    if (s->reg_accumulate && !sub)
        return fail(HIPCLAW_EINVAL, "hipclaw_autotune: a subcycling state needs its levels set%s");
    *ntrials = 0;
    if (s->npatch == 0) return HIPCLAW_OK;
    if (sub)
        if (int rc = subcycle_ready(s)) return rc;
    auto trial_steps = [&](int n) { return sub ? subcycle_launches(s, dt, n) : advance_launches(s, dt, n, false); };
    // the trials step through both slabs: the state is copied, not kept in its slab
    if (int rc = copy_to_save(s, s->cur)) return rc;
    s->save_lazy = false;
    s->saved = true;
    hipEvent_t ev[2] = {nullptr, nullptr};
    HIP_TRY(hipEventCreate(&ev[0]));
    HIP_TRY(hipEventCreate(&ev[1]));
    int best_tile = s->tile, best_layout = s->ghost_layout, best_waves = s->waves, n = 0, rc = HIPCLAW_OK;
    double best_ms = 0.0;
    static const int candidates[3] = {8, 16, 32};
    for (int t : candidates) {
        if (!tile_allowed(s, t)) continue;
        const int nlay = s->fused_ok ? 3 : 2;
        for (int trial = 0; trial < (t == 16 ? 3 * nlay : nlay) && !rc; trial++) {
            const int lay = trial % nlay, w = t == 16 ? 2 + trial / nlay : 0;
            if ((rc = set_tile(s, t)) || (rc = hipclaw_set_ghost_layout(s, lay)) || (rc = hipclaw_set_waves(s, w)))
                break;
            // one untimed step (first use of these kernels), then the timed ones
            if ((rc = trial_steps(1))) break;
            hipError_t e = hipEventRecord(ev[0], s->st);
            if (e == hipSuccess && (rc = trial_steps(nsteps))) break;
            if (e == hipSuccess) e = hipEventRecord(ev[1], s->st);
            if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
            float elapsed = 0.0f;
            if (e == hipSuccess) e = hipEventElapsedTime(&elapsed, ev[0], ev[1]);
            if (e != hipSuccess) { rc = hipfail(e, "hipclaw_autotune timing"); break; }
            if ((rc = hipclaw_restore_step(s))) break;
            tiles[n] = t;
            waves[n] = effective_waves(s);
            layouts[n] = lay;
            ms[n] = (double) elapsed / nsteps;
            if (n == 0 || ms[n] < best_ms) { best_tile = t; best_layout = lay; best_waves = w; best_ms = ms[n]; }
            n++;
        }
        if (rc) break;
    }
    hipEventDestroy(ev[0]);
    hipEventDestroy(ev[1]);
    if (rc) {
        hipclaw_restore_step(s);
        return rc;
    }
    *ntrials = n;
    if (int r = set_tile(s, best_tile)) return r;
    if (int r = hipclaw_set_waves(s, best_waves)) return r;
    return hipclaw_set_ghost_layout(s, best_layout);
}

int hipclaw_set_dt_scale(hipclaw_state_t* s, const double* scale) {
    if (!s) return fail(HIPCLAW_EINVAL, "hipclaw_set_dt_scale: bad arguments%s");
    if (!scale) {
        HIP_TRY(hipStreamSynchronize(s->st));
        hipFree(s->dtscale);
        s->dtscale = nullptr;
    } else {
        for (int p = 0; p < s->npatch; p++)
            if (!(scale[p] > 0.0) || !std::isfinite(scale[p]))
                return fail(HIPCLAW_EINVAL, "dt scale must be positive and finite%s");
        if (!s->dtscale) HIP_TRY(hipMalloc((void**) &s->dtscale, sizeof(double) * (size_t) s->max_patches));
        // ordered after the steps already queued on the (non-blocking) stream
        HIP_TRY(hipMemcpyAsync(s->dtscale, scale, sizeof(double) * (size_t) s->npatch,
                               hipMemcpyHostToDevice, s->st));
        HIP_TRY(hipStreamSynchronize(s->st));
    }
    // a captured graph holds the old step arguments
    if (s->gexec) { hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { hipGraphDestroy(s->graph); s->graph = nullptr; }
    s->graph_steps = 0;
    return HIPCLAW_OK;
}

int hipclaw_vtk_values(hipclaw_state_t* s, int first, int count, float* out) {
    if (!s || first < 0 || count < 0 || first + count > s->npatch || (count > 0 && !out))
        return fail(HIPCLAW_EINVAL, "hipclaw_vtk_values: bad arguments%s");
    if (count == 0) return HIPCLAW_OK;
    const size_t n = (size_t) count * s->cfg.mx * s->cfg.my * s->cfg.meqn;
    float* dev = nullptr;
    HIP_TRY(hipMalloc((void**) &dev, sizeof(float) * n));
    hipError_t e = (hipError_t) hipclaw::launch_vtk_cell_values(s->q[s->cur], first, count, s->pstride,
                                                                s->cfg.mx, s->cfg.my, s->cfg.mbc,
                                                                s->cfg.meqn, dev, s->st);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dev, sizeof(float) * n, hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    hipFree(dev);
    if (e != hipSuccess) return hipfail(e, "vtk_cell_values");
    return HIPCLAW_OK;
}

int hipclaw_get_registers(hipclaw_state_t* s, int first, int count, double* out) {
    if (!s || first < 0 || count < 0 || first + count > s->npatch || (count > 0 && !out))
        return fail(HIPCLAW_EINVAL, "hipclaw_get_registers: bad arguments%s");
    if (!s->reg) return fail(HIPCLAW_EINVAL, "the state was created without time_sync%s");
    HIP_TRY(hipStreamSynchronize(s->st));
    HIP_TRY(hipMemcpy(out, s->reg + (long long) first * s->rstride,
                      sizeof(double) * s->rstride * (size_t) count, hipMemcpyDeviceToHost));
    return HIPCLAW_OK;
}

int hipclaw_timeinterp(hipclaw_state_t* s, const int* slots, int nslots, double alpha) {
    if (!s || nslots < 0 || nslots > s->max_patches || (nslots > 0 && !slots))
        return fail(HIPCLAW_EINVAL, "hipclaw_timeinterp: bad arguments%s");
    if (!s->qti) return fail(HIPCLAW_EINVAL, "the state was created without subcycle%s");
    for (int k = 0; k < nslots; k++)
        if (slots[k] < 0 || slots[k] >= s->npatch) return fail(HIPCLAW_EINVAL, "slot out of range%s");
    if (nslots == 0) return HIPCLAW_OK;
    HIP_TRY(hipMemcpyAsync(s->batch_slots, slots, sizeof(int) * nslots, hipMemcpyHostToDevice, s->st));
    int e = hipclaw::launch_timeinterp(s->q[s->cur], s->qlast, s->qti, s->batch_slots, nslots,
                                       s->pstride, s->cfg.mx, s->cfg.my, s->cfg.mbc, s->cfg.meqn,
                                       alpha, s->st);
    if (e) return hipfail((hipError_t) e, "timeinterp launch");
    HIP_TRY(hipStreamSynchronize(s->st));
    return HIPCLAW_OK;
}

static hipclaw::StepArgs step_args(hipclaw_state_t* s, const double* src, double* dst,
                                   const int* slots, int n, double dt) {
    hipclaw::StepArgs a;
    a.qsrc = src;
    a.qdst = dst;
    a.aux = s->aux;
    a.work = slots;
    a.dxs = s->dxs;
    a.dys = s->dys;
    a.cfl = s->cfl;
    a.pstride = s->pstride;
    a.astride = s->astride;
    a.nwork = n;
    a.mx = s->cfg.mx; a.my = s->cfg.my; a.mbc = s->cfg.mbc;
    a.ntx = s->ntx; a.nty = s->nty;
    set_method(a, &s->cfg);
    a.dt = dt;
    a.reg = s->reg;
    a.rstride = s->rstride;
    a.reg_accumulate = s->reg_accumulate;
    a.dtscale = s->dtscale;
    return a;
}

// ev0 / ev1: optional events bound to the kernel's own start and end.
static int launch_step(hipclaw_state_t* s, const hipclaw::StepArgs& a, hipEvent_t ev0 = nullptr,
                       hipEvent_t ev1 = nullptr) {
    const hipclaw::StepLaunch L(s->st, ev0, ev1);
    int e = launch_solver_step(s->cfg.solver, s->cfg.precision, s->tile, s->waves, a, L);
    if (e) return hipfail((hipError_t) e, "step_tiles launch");
    return HIPCLAW_OK;
}

static double read_cfl(hipclaw_state_t* s, int* rc) {
    unsigned long long bits = 0;
    hipError_t e = hipMemcpyAsync(&bits, s->cfl, sizeof(bits), hipMemcpyDeviceToHost, s->st);
    if (e == hipSuccess) e = hipStreamSynchronize(s->st);
    if (e != hipSuccess) { *rc = hipfail(e, "reading the CFL number"); return 0.0; }
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    *rc = HIPCLAW_OK;
    return v;
}

int hipclaw_step_batch(hipclaw_state_t* s, const int* slots, int nslots, double t, double dt,
                       double* maxcfl) {
    (void) t;
    if (!s || nslots < 0 || nslots > s->max_patches || (nslots > 0 && !slots) || !(dt > 0.0))
        return fail(HIPCLAW_EINVAL, "hipclaw_step_batch: bad arguments%s");
    for (int k = 0; k < nslots; k++)
        if (slots[k] < 0 || slots[k] >= s->npatch) return fail(HIPCLAW_EINVAL, "slot out of range%s");
    if (int rc = guard_save(s)) return rc;   // the batch writes both slabs
    if (maxcfl) *maxcfl = 0.0;
    if (nslots == 0) return HIPCLAW_OK;
    HIP_TRY(hipMemcpyAsync(s->batch_slots, slots, sizeof(int) * nslots, hipMemcpyHostToDevice, s->st));
    if (int e = hipclaw::launch_reset_scalar(s->cfl, s->st)) return hipfail((hipError_t) e, "reset");
    if (s->qlast) {  // fclaw2d_clawpatch_save_current_step before every update
        int e = hipclaw::launch_copy_patches(s->q[s->cur], s->qlast, s->batch_slots, nslots,
                                             s->pstride, s->st);
        if (e) return hipfail((hipError_t) e, "copy_patches launch");
    }
    double* src = s->q[s->cur];
    double* dst = s->inplace ? src : s->q[1 - s->cur];
    if (int rc = launch_step(s, step_args(s, src, dst, s->batch_slots, nslots, dt))) return rc;
    if (!s->inplace) {
        int e = hipclaw::launch_copy_interior(dst, src, s->batch_slots, nslots, s->pstride,
                                              s->cfg.mx, s->cfg.my, s->cfg.mbc, s->cfg.meqn, s->st);
        if (e) return hipfail((hipError_t) e, "copy_interior launch");
    }
    int rc;
    const double c = read_cfl(s, &rc);
    if (rc) return rc;
    if (maxcfl) *maxcfl = c;
    return HIPCLAW_OK;
}

int hipclaw_step_batch_noalloc(const hipclaw_config_t* cfg, int npatch, const double* q_in,
                               double* q_out, const double* aux, const double* dx,
                               const double* dy, const int* slots, int nslots, double dt,
                               void* stream, double* maxcfl) {
    if (!cfg || npatch < 0 || nslots < 0 || (nslots > 0 && (!slots || !q_in || !q_out || !dx || !dy)) ||
        !(dt > 0.0))
        return fail(HIPCLAW_EINVAL, "hipclaw_step_batch_noalloc: bad arguments%s");
    if (nslots > HIPCLAW_NOALLOC_MAX_SLOTS)
        return fail(HIPCLAW_EINVAL, "hipclaw_step_batch_noalloc: more slots than the scratch holds "
                                    "(HIPCLAW_NOALLOC_MAX_SLOTS)%s");
    int tile = 0;
    if (int rc = check_config(cfg, &tile)) return rc;
    if ((cfg->maux > 0) != (aux != nullptr))
        return fail(HIPCLAW_EINVAL, "hipclaw_step_batch_noalloc: aux must be given exactly when maux > 0%s");
    const int ntx = tiles_along(cfg->mx, tile), nty = tiles_along(cfg->my, tile);
    if (q_in == q_out && (ntx > 1 || nty > 1))
        return fail(HIPCLAW_EINVAL, "hipclaw_step_batch_noalloc: q_in and q_out must differ when a "
                                    "tile does not cover a patch%s");
    for (int k = 0; k < nslots; k++)
        if (slots[k] < 0 || slots[k] >= npatch) return fail(HIPCLAW_EINVAL, "slot out of range%s");
    if (maxcfl) *maxcfl = 0.0;
    if (nslots == 0) return HIPCLAW_OK;
    int rc;
    unsigned char* base = static_cast<unsigned char*>(scratch_ptr(&rc));
    if (!base) return rc;
    unsigned long long* dcfl = reinterpret_cast<unsigned long long*>(base);
    int* dslots = reinterpret_cast<int*>(base + NOALLOC_CFL_BYTES);
    hipStream_t st = static_cast<hipStream_t>(stream);
    HIP_TRY(hipMemcpyAsync(dslots, slots, sizeof(int) * (size_t) nslots, hipMemcpyHostToDevice, st));
    if (int e = hipclaw::launch_reset_scalar(dcfl, st)) return hipfail((hipError_t) e, "reset");
    hipclaw::StepArgs a;
    a.qsrc = q_in;
    a.qdst = q_out;
    a.aux = aux;
    a.work = dslots;
    a.dxs = dx;
    a.dys = dy;
    a.cfl = dcfl;
    const long long cells = (long long) (cfg->mx + 2 * cfg->mbc) * (cfg->my + 2 * cfg->mbc);
    a.pstride = cells * cfg->meqn;
    a.astride = cells * cfg->maux;
    a.nwork = nslots;
    a.mx = cfg->mx; a.my = cfg->my; a.mbc = cfg->mbc;
    a.ntx = ntx; a.nty = nty;
    set_method(a, cfg);
    a.dt = dt;
    a.reg = nullptr;
    a.rstride = 0;
    a.reg_accumulate = 0;
    a.dtscale = nullptr;
    const hipclaw::StepLaunch L(st);
    int e = launch_solver_step(cfg->solver, cfg->precision, tile, 0, a, L);
    if (e) return hipfail((hipError_t) e, "step_tiles launch");
    unsigned long long bits = 0;
    HIP_TRY(hipMemcpyAsync(&bits, dcfl, sizeof(bits), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    if (maxcfl) *maxcfl = v;
    return HIPCLAW_OK;
}

static int advance_launches(hipclaw_state_t* s, double dt, int nsteps, bool timed) {
    if (timed) {
        const size_t need = 2 * (size_t) nsteps;
        while (s->kev.size() < need) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            s->kev.push_back(e);
        }
    }
    if (int e = hipclaw::launch_reset_scalar(s->cfl, s->st)) return hipfail((hipError_t) e, "reset");
    if (s->ghost_layout == 2 && s->fused_ok && !s->reg_accumulate) {
        // two launches per step: the update (staging the corrected edge cells of
        // the last synchronisation) and the fused ghost fill + time sync; the
        // corrections of the last step are written into the slab at the end
        for (int n = 0; n < nsteps; n++) {
            double* src = s->q[s->cur];
            double* dst = s->inplace ? src : s->q[1 - s->cur];
            hipclaw::StepArgs a = step_args(s, src, dst, s->all_slots, s->npatch, dt);
            if (s->fsync && n > 0) {
                a.qcorr = s->qcorr;
                a.cmask = s->cmask;
                a.cstride = s->cstride;
            }
            // Pulling reads the neighbours' cells of the slab the update reads, so
            // only when it writes the other one, and not in the first step: the
            // slab then holds the last advance's synchronised interior cells, while
            // its ghost cells are the copies made before that synchronisation.
            const bool pull = s->pull_ok && !s->inplace;
            if (pull && n > 0) a.pull = s->pull;
            if (int rc = timed ? launch_step(s, a, s->kev[2 * n], s->kev[2 * n + 1]) : launch_step(s, a)) return rc;
            if (!s->inplace) s->cur = 1 - s->cur;
            if (int rc = ghost_flat_async(s, !pull || n == nsteps - 1, s->fsync)) return rc;
        }
        if (s->fsync && nsteps > 0) {
            int e = hipclaw::launch_apply_qcorr(s->q[s->cur], s->qcorr, s->cmask, s->qslots, s->nqslots,
                                                s->pstride, s->cstride, s->cfg.mx, s->cfg.my, s->cfg.mbc,
                                                s->cfg.meqn, s->st);
            if (e) return hipfail((hipError_t) e, "apply_qcorr launch");
        }
        return HIPCLAW_OK;
    }
    for (int n = 0; n < nsteps; n++) {
        double* src = s->q[s->cur];
        double* dst = s->inplace ? src : s->q[1 - s->cur];
        const hipclaw::StepArgs a = step_args(s, src, dst, s->all_slots, s->npatch, dt);
        if (int rc = timed ? launch_step(s, a, s->kev[2 * n], s->kev[2 * n + 1]) : launch_step(s, a)) return rc;
        if (!s->inplace) s->cur = 1 - s->cur;
        if (int rc = ghost_fill_async(s, 0)) return rc;
        if (int rc = time_sync_async(s, 0)) return rc;
    }
    return HIPCLAW_OK;
}

// ---- subcycled advance on the device ----
//
// fclaw2d_advance_all_levels with subcycle = T (src/fclaw2d_advance.c:54-246):
// level L takes 2^(L - lmin) steps of dt / 2^(L - lmin) per coarse step; after
// every finest step that is not the last, the ghost cells of the levels above
// the coarsest level still between steps are filled from time-interpolated
// copies (ghost program 1 + ti - lmin) and, with time sync, those levels are
// synchronised.  Every launch goes on the library's stream in the order the
// reference's host loop makes its calls, with no host synchronisation: the
// Courant numbers of all substeps meet in the one device maximum, so the
// whole coarse step can be captured in a graph.

// One level's update: the state before it (fclaw2d_clawpatch_save_current_step)
// copied whole into qlast, then the step reading that copy and writing the
// new interiors straight into the current slab, which the other levels keep
// reading (the same values as stepping the slab itself: qlast holds its
// every cell).
static int subcycle_update(hipclaw_state_t* s, int level, double dt) {
    const int l = level - s->lmin, n = s->loff[l + 1] - s->loff[l];
    if (n == 0) return HIPCLAW_OK;
    const int* slots = s->lslots + s->loff[l];
    int e = hipclaw::launch_copy_patches(s->q[s->cur], s->qlast, slots, n, s->pstride, s->st);
    if (e) return hipfail((hipError_t) e, "copy_patches launch");
    return launch_step(s, step_args(s, s->qlast, s->q[s->cur], slots, n, dt));
}

// The time-interpolated copies of a level (fclaw2d_timeinterp).
static int subcycle_timeinterp(hipclaw_state_t* s, int level, double alpha) {
    const int l = level - s->lmin, n = s->loff[l + 1] - s->loff[l];
    if (n == 0) return HIPCLAW_OK;
    int e = hipclaw::launch_timeinterp(s->q[s->cur], s->qlast, s->qti, s->lslots + s->loff[l], n,
                                       s->pstride, s->cfg.mx, s->cfg.my, s->cfg.mbc, s->cfg.meqn,
                                       alpha, s->st);
    if (e) return hipfail((hipError_t) e, "timeinterp launch");
    return HIPCLAW_OK;
}

namespace {
struct LevelCounter {
    int last_step = 0, step_inc = 1, total_steps = 1;
    double dt_step = 0.0;
};
}  // namespace

// advance_level (src/fclaw2d_advance.c:112-150, 152-170): the level's step,
// then, when the coarser level is due, its step and its time-interpolated
// copies at the finer level's new time.
static int subcycle_level(hipclaw_state_t* s, int level, int curr_fine_step, LevelCounter* tc) {
    LevelCounter& c = tc[level - s->lmin];
    if (int rc = subcycle_update(s, level, c.dt_step)) return rc;
    c.last_step += c.step_inc;
    if (level > s->lmin) {
        LevelCounter& cc = tc[level - 1 - s->lmin];
        if (cc.last_step == curr_fine_step) {
            if (int rc = subcycle_level(s, level - 1, cc.last_step, tc)) return rc;
            const double alpha = ((double) (c.last_step % cc.step_inc)) / cc.step_inc;
            if (int rc = subcycle_timeinterp(s, level - 1, alpha)) return rc;
        }
    }
    return HIPCLAW_OK;
}

static int subcycle_launches(hipclaw_state_t* s, double dt, int ncoarse) {
    const int lmin = s->lmin, lmax = s->lmax, nl = lmax - lmin + 1;
    if (int e = hipclaw::launch_reset_scalar(s->cfl, s->st)) return hipfail((hipError_t) e, "reset");
    for (int c = 0; c < ncoarse; c++) {
        LevelCounter tc[HIPCLAW_MAX_PROGRAMS];
        for (int l = 0; l < nl; l++) {   // initialize_timestep_counters (:54-110)
            tc[l].dt_step = dt / (double) (1 << l);
            tc[l].total_steps = 1 << l;
            tc[l].step_inc = 1 << (lmax - lmin - l);
        }
        const int total = tc[nl - 1].total_steps;
        for (int nf = 0; nf < total; nf++) {
            if (int rc = subcycle_level(s, lmax, nf, tc)) return rc;
            const int last_step = tc[nl - 1].last_step;
            if (last_step < total) {
                int ti = lmax - 1;
                while (last_step % tc[ti - lmin].step_inc == 0) ti--;
                const int k = 1 + ti - lmin;   // levels ti + 1 .. lmax, ti time-interpolated
                if (int rc = ghost_fill_async(s, k)) return rc;
                if (s->reg)
                    if (int rc = time_sync_async(s, k)) return rc;
            }
        }
        if (int rc = ghost_fill_async(s, 0)) return rc;
        if (s->reg)
            if (int rc = time_sync_async(s, 0)) return rc;
    }
    return HIPCLAW_OK;
}

// What the subcycled advance needs: subcycle state, levels, and the programs
// of every level range.
static int subcycle_ready(hipclaw_state_t* s) {
    if (!s->qti || !s->qlast) return fail(HIPCLAW_EINVAL, "the state was created without subcycle%s");
    if (s->lmax < s->lmin || (int) s->loff.size() != s->lmax - s->lmin + 2)
        return fail(HIPCLAW_EINVAL, "no levels set (hipclaw_set_levels)%s");
    // program 0: every level; program k = 1 + ti - lmin: levels ti + 1 .. lmax
    // with level ti time-interpolated (ghost fill), levels from ti + 1 (sync)
    for (int k = 0; k <= s->lmax - s->lmin; k++) {
        if (!s->gp[k].set) return fail(HIPCLAW_EINVAL, "a ghost program of the subcycled schedule is missing%s");
        if (s->reg && !s->sp[k].set)
            return fail(HIPCLAW_EINVAL, "a time-sync program of the subcycled schedule is missing%s");
    }
    if (s->reg && s->sp[0].nresets == 0 && s->sp[0].ngroups > 0)
        return fail(HIPCLAW_EINVAL, "with subcycle, the full sync program needs its register resets%s");
    return ensure_layouts(s);
}

int hipclaw_set_levels(hipclaw_state_t* s, const int* level, int npatch) {
    if (!s || npatch != s->npatch || (npatch > 0 && !level))
        return fail(HIPCLAW_EINVAL, "hipclaw_set_levels: one level per patch%s");
    drop_graph(s);
    if (npatch == 0) { s->lmax = s->lmin - 1; return HIPCLAW_OK; }
    int lo = level[0], hi = level[0];
    for (int p = 0; p < npatch; p++) {
        if (level[p] < 0 || level[p] > 30) return fail(HIPCLAW_EINVAL, "hipclaw_set_levels: level out of range%s");
        lo = std::min(lo, level[p]);
        hi = std::max(hi, level[p]);
    }
    if (hi - lo + 1 > HIPCLAW_MAX_PROGRAMS || hi - lo > 20)
        return fail(HIPCLAW_EINVAL, "hipclaw_set_levels: more levels than programs%s");
    std::vector<int> off(hi - lo + 2, 0), slots(npatch);
    for (int p = 0; p < npatch; p++) off[level[p] - lo + 1]++;
    for (int l = 0; l <= hi - lo; l++) off[l + 1] += off[l];
    std::vector<int> fill(off.begin(), off.end() - 1);
    for (int p = 0; p < npatch; p++) slots[fill[level[p] - lo]++] = p;
    if (int rc = upload_array(s, s->lslots, s->lslots_cap, slots.data(), npatch, "level slots")) return rc;
    s->loff = off;
    s->lmin = lo;
    s->lmax = hi;
    return HIPCLAW_OK;
}

int hipclaw_advance_subcycle_async(hipclaw_state_t* s, double dt, int ncoarse) {
    if (!s || ncoarse < 0 || !(dt > 0.0)) return fail(HIPCLAW_EINVAL, "hipclaw_advance_subcycle: bad arguments%s");
    if (int rc = subcycle_ready(s)) return rc;
    if (int rc = guard_save(s)) return rc;
    HIP_TRY(hipEventRecord(s->region[0], s->st));
    if (int rc = subcycle_launches(s, dt, ncoarse)) return rc;
    HIP_TRY(hipEventRecord(s->region[1], s->st));
    s->timed_steps = 0;
    return HIPCLAW_OK;
}

int hipclaw_advance_subcycle(hipclaw_state_t* s, double dt, int ncoarse, double* maxcfl) {
    if (int rc = hipclaw_advance_subcycle_async(s, dt, ncoarse)) return rc;
    int rc;
    const double c = read_cfl(s, &rc);
    if (rc) return rc;
    if (maxcfl) *maxcfl = c;
    return HIPCLAW_OK;
}

int hipclaw_capture_subcycle_graph(hipclaw_state_t* s, double dt, int ncoarse) {
    if (!s || ncoarse <= 0 || !(dt > 0.0))
        return fail(HIPCLAW_EINVAL, "hipclaw_capture_subcycle_graph: bad arguments%s");
    if (s->gexec) { hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { hipGraphDestroy(s->graph); s->graph = nullptr; }
    if (int rc = subcycle_ready(s)) return rc;
    if (int rc = guard_save(s)) return rc;   // (replays write the slabs)
    HIP_TRY(hipStreamSynchronize(s->st));
    s->graph_cur = s->cur;   // the subcycled steps write the current slab in place
    HIP_TRY(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    int rc = subcycle_launches(s, dt, ncoarse);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->st, &g);
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return hipfail(e, "hipStreamEndCapture");
    s->graph = g;
    HIP_TRY(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
    s->graph_steps = ncoarse;
    return HIPCLAW_OK;
}

int hipclaw_advance_async(hipclaw_state_t* s, double dt, int nsteps) {
    if (!s || nsteps < 0 || !(dt > 0.0)) return fail(HIPCLAW_EINVAL, "hipclaw_advance: bad arguments%s");
    if (s->reg_accumulate && s->sp[0].nresets == 0 && s->sp[0].ngroups > 0)
        return fail(HIPCLAW_EINVAL, "with subcycle, the full sync program needs its register resets%s");
    if (!(s->ghost_layout == 2 && s->fused_ok && !s->reg_accumulate))
        if (int rc = ensure_layouts(s)) return rc;
    if (int rc = guard_save(s, nsteps == 1)) return rc;
    HIP_TRY(hipEventRecord(s->region[0], s->st));
    if (int rc = advance_launches(s, dt, nsteps, s->time_kernels)) return rc;
    HIP_TRY(hipEventRecord(s->region[1], s->st));
    s->timed_steps = s->time_kernels ? nsteps : 0;
    return HIPCLAW_OK;
}

int hipclaw_sync(hipclaw_state_t* s) {
    if (!s) return fail(HIPCLAW_EINVAL, "null state%s");
    HIP_TRY(hipStreamSynchronize(s->st));
    return HIPCLAW_OK;
}

int hipclaw_advance(hipclaw_state_t* s, double dt, int nsteps, double* maxcfl) {
    if (int rc = hipclaw_advance_async(s, dt, nsteps)) return rc;
    int rc;
    const double c = read_cfl(s, &rc);
    if (rc) return rc;
    if (maxcfl) *maxcfl = c;
    return HIPCLAW_OK;
}

int hipclaw_capture_graph(hipclaw_state_t* s, double dt, int nsteps) {
    if (!s || nsteps <= 0 || !(dt > 0.0)) return fail(HIPCLAW_EINVAL, "hipclaw_capture_graph: bad arguments%s");
    if (!s->inplace && (nsteps % 2) != 0)
        return fail(HIPCLAW_EINVAL, "ping-pong states need an even number of captured steps%s");
    if (s->gexec) { hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { hipGraphDestroy(s->graph); s->graph = nullptr; }
    if (int rc = guard_save(s)) return rc;   // (replays write both slabs)
    HIP_TRY(hipStreamSynchronize(s->st));
    if (!(s->ghost_layout == 2 && s->fused_ok && !s->reg_accumulate))
        if (int rc = ensure_layouts(s)) return rc;
    const int cur0 = s->cur;
    s->graph_cur = cur0;
    HIP_TRY(hipStreamBeginCapture(s->st, hipStreamCaptureModeThreadLocal));
    int rc = advance_launches(s, dt, nsteps, false);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(s->st, &g);
    s->cur = cur0;  // capture only recorded the launches
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return hipfail(e, "hipStreamEndCapture");
    s->graph = g;
    HIP_TRY(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
    s->graph_steps = nsteps;
    return HIPCLAW_OK;
}

int hipclaw_replay_graph(hipclaw_state_t* s) {
    if (!s || !s->gexec) return fail(HIPCLAW_EINVAL, "no captured graph%s");
    if (s->cur != s->graph_cur)
        return fail(HIPCLAW_EINVAL, "the captured graph starts from the other slab (an odd number of "
                                    "steps or a restored step since the capture)%s");
    if (int rc = guard_save(s)) return rc;
    HIP_TRY(hipEventRecord(s->region[0], s->st));
    HIP_TRY(hipGraphLaunch(s->gexec, s->st));
    HIP_TRY(hipEventRecord(s->region[1], s->st));
    s->timed_steps = 0;
    // an even number of ping-pong steps leaves the current slab unchanged
    return HIPCLAW_OK;
}

int hipclaw_timing(hipclaw_state_t* s, int nsteps, double* total_ms, double* step_kernel_ms) {
    if (!s) return fail(HIPCLAW_EINVAL, "null state%s");
    HIP_TRY(hipStreamSynchronize(s->st));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, s->region[0], s->region[1]));
    if (total_ms) *total_ms = ms;
    if (step_kernel_ms) {
        *step_kernel_ms = 0.0;
        const int n = std::min(nsteps, s->timed_steps);
        double sum = 0.0;
        for (int k = 0; k < n; k++) {
            float m = 0.f;
            HIP_TRY(hipEventElapsedTime(&m, s->kev[2 * k], s->kev[2 * k + 1]));
            sum += m;
        }
        if (n > 0) *step_kernel_ms = sum / n;
    }
    return HIPCLAW_OK;
}

static void two_sum_host(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

int hipclaw_mass(hipclaw_state_t* s, double* sum_hi, double* sum_lo) {
    if (!s || !sum_hi) return fail(HIPCLAW_EINVAL, "hipclaw_mass: bad arguments%s");
    const int meqn = s->cfg.meqn;
    int e = hipclaw::launch_mass(s->q[s->cur], s->all_slots, s->npatch, s->dxs, s->dys, s->pstride,
                                 s->cfg.mx, s->cfg.my, s->cfg.mbc, meqn, s->mass_buf, s->st);
    if (e) return hipfail((hipError_t) e, "patch_mass launch");
    std::vector<double> buf(2 * (size_t) meqn * s->npatch);
    if (!buf.empty())
        HIP_TRY(hipMemcpyAsync(buf.data(), s->mass_buf, sizeof(double) * buf.size(), hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    for (int m = 0; m < meqn; m++) {
        double hi = 0.0, lo = 0.0;
        for (int p = 0; p < s->npatch; p++) {
            double t, err;
            two_sum_host(hi, buf[((size_t) p * meqn + m) * 2], t, err);
            hi = t;
            lo += err + buf[((size_t) p * meqn + m) * 2 + 1];
        }
        double t, err;
        two_sum_host(hi, lo, t, err);
        sum_hi[m] = t;
        if (sum_lo) sum_lo[m] = err;
    }
    return HIPCLAW_OK;
}

// s->iwork with room for n ints.
static int ensure_iwork(hipclaw_state_t* s, size_t n) {
    if (n <= s->iwork_cap) return HIPCLAW_OK;
    HIP_TRY(hipStreamSynchronize(s->st));
    hipFree(s->iwork);
    s->iwork = nullptr;
    s->iwork_cap = 0;
    const size_t cap = n + n / 4;
    if (hipMalloc((void**) &s->iwork, sizeof(int) * cap) != hipSuccess)
        return fail(HIPCLAW_ENOMEM, "hipMalloc of the regrid workspace failed%s");
    s->iwork_cap = cap;
    return HIPCLAW_OK;
}

int hipclaw_tag_patches(hipclaw_state_t* s, int ivar, double refine_threshold,
                        double coarsen_threshold, int* exceeds_refine, int* exceeds_coarsen) {
    if (!s || ivar < 0 || ivar >= s->cfg.meqn || !exceeds_refine || !exceeds_coarsen)
        return fail(HIPCLAW_EINVAL, "hipclaw_tag_patches: bad arguments%s");
    const int n = s->npatch;
    if (n == 0) return HIPCLAW_OK;
    if (int rc = ensure_iwork(s, 2 * (size_t) n)) return rc;
    int* dev = s->iwork;
    int e = hipclaw::launch_tag_patches(s->q[s->cur], s->all_slots, n, s->dxs, s->dys, s->pstride,
                                        s->cfg.mx, s->cfg.my, s->cfg.mbc, ivar, refine_threshold,
                                        coarsen_threshold, dev, s->st);
    std::vector<int> out(2 * (size_t) n);
    hipError_t he = e ? (hipError_t) e
                      : hipMemcpyAsync(out.data(), dev, sizeof(int) * 2 * (size_t) n, hipMemcpyDeviceToHost, s->st);
    if (he == hipSuccess) he = hipStreamSynchronize(s->st);
    if (he != hipSuccess) return hipfail(he, "tag_patches");
    for (int p = 0; p < n; p++) {
        exceeds_refine[p] = out[2 * p];
        exceeds_coarsen[p] = out[2 * p + 1];
    }
    return HIPCLAW_OK;
}

// Grow every per-patch device buffer to `cap` patches.  q is replaced by the
// caller; everything else starts at zero (registers, saved and interpolated
// states, aux) or is rebuilt (slot list).
static int grow_patch_buffers(hipclaw_state_t* s, int cap) {
    const size_t P = (size_t) cap;
    auto realloc0 = [&](double*& p, size_t elems) -> int {
        hipFree(p);
        p = nullptr;
        if (hipMalloc((void**) &p, sizeof(double) * elems) != hipSuccess) return HIPCLAW_ENOMEM;
        return hipMemset(p, 0, sizeof(double) * elems) == hipSuccess ? HIPCLAW_OK : HIPCLAW_EDEVICE;
    };
    int rc = HIPCLAW_OK;
    if (s->cfg.maux > 0 && !rc) rc = realloc0(s->aux, s->astride * P);
    if (!rc) rc = realloc0(s->dxs, P);
    if (!rc) rc = realloc0(s->dys, P);
    if (!rc) rc = realloc0(s->mass_buf, 2 * (size_t) s->cfg.meqn * P);
    if (!rc && s->qlast) rc = realloc0(s->qlast, s->pstride * P);
    if (!rc && s->qti) rc = realloc0(s->qti, s->pstride * P);
    if (!rc && s->reg) rc = realloc0(s->reg, s->rstride * P);
    if (rc) return fail(rc, "hipclaw_regrid: device allocation failed%s");
    hipFree(s->all_slots);
    hipFree(s->batch_slots);
    s->all_slots = s->batch_slots = nullptr;
    if (hipMalloc((void**) &s->all_slots, sizeof(int) * P) != hipSuccess ||
        hipMalloc((void**) &s->batch_slots, sizeof(int) * P) != hipSuccess)
        return fail(HIPCLAW_ENOMEM, "hipclaw_regrid: device allocation failed%s");
    std::vector<int> iota(P);
    for (int p = 0; p < cap; p++) iota[p] = p;
    HIP_TRY(hipMemcpy(s->all_slots, iota.data(), sizeof(int) * P, hipMemcpyHostToDevice));
    s->max_patches = cap;
    return HIPCLAW_OK;
}

int hipclaw_regrid(hipclaw_state_t* s, int npatch_new, const int* kind, const int* src,
                   const double* dx, const double* dy) {
    if (!s || npatch_new <= 0 || !kind || !src || !dx || !dy)
        return fail(HIPCLAW_EINVAL, "hipclaw_regrid: bad arguments%s");
    const int nold = s->npatch;
    std::vector<int> ops(8 * (size_t) npatch_new, 0);
    for (int p = 0; p < npatch_new; p++) {
        const int k = kind[p];
        const int* sp = src + 4 * (size_t) p;
        int* op = ops.data() + 8 * (size_t) p;
        auto ok = [&](int v) { return v >= 0 && v < nold; };
        if (!(dx[p] > 0.0) || !(dy[p] > 0.0)) return fail(HIPCLAW_EINVAL, "non-positive mesh width%s");
        if (k == hipclaw::REGRID_SAME ? !ok(sp[0])
            : k == hipclaw::REGRID_CHILD ? !(ok(sp[0]) && sp[1] >= 0 && sp[1] <= 3)
            : k == hipclaw::REGRID_PARENT ? !(ok(sp[0]) && ok(sp[1]) && ok(sp[2]) && ok(sp[3]))
            : true)
            return fail(HIPCLAW_EINVAL, "regrid source with bad kind or slot%s");
        if (k != hipclaw::REGRID_SAME && s->cfg.mx != s->cfg.my)
            return fail(HIPCLAW_EINVAL, "refining or coarsening a patch requires mx = my%s");
        op[0] = k;
        op[1] = p;
        for (int g = 0; g < 4; g++) op[2 + g] = sp[g];
    }
    // destination slab: the spare ping-pong slab when the new patches fit,
    // otherwise a larger allocation (with room to grow) replacing both slabs
    const bool reuse = !s->inplace && npatch_new <= s->max_patches;
    const int cap = reuse ? s->max_patches
                          : std::max(npatch_new, std::max(s->max_patches, npatch_new + npatch_new / 4));
    double* qnew = nullptr;
    if (int rc = ensure_iwork(s, ops.size())) return rc;
    int* dops = s->iwork;
    HIP_TRY(hipStreamSynchronize(s->st));
    if (reuse) qnew = s->q[1 - s->cur];
    else if (hipMalloc((void**) &qnew, sizeof(double) * s->pstride * (size_t) cap) != hipSuccess)
        return fail(HIPCLAW_ENOMEM, "hipclaw_regrid: device allocation failed%s");
    // (regrid_remap zeroes the ghost cells of new children and parents)
    hipError_t he = hipMemcpyAsync(dops, ops.data(), sizeof(int) * ops.size(), hipMemcpyHostToDevice, s->st);
    if (he == hipSuccess)
        he = (hipError_t) hipclaw::launch_regrid_remap(s->q[s->cur], qnew, dops, npatch_new, s->pstride,
                                                       s->cfg.mx, s->cfg.my, s->cfg.mbc, s->cfg.meqn, s->st);
    if (he == hipSuccess) he = hipStreamSynchronize(s->st);
    if (he != hipSuccess) {
        if (!reuse) hipFree(qnew);
        return hipfail(he, "regrid_remap");
    }
    if (reuse) {
        s->cur = 1 - s->cur;
    } else {
        hipFree(s->q[0]);
        hipFree(s->q[1]);
        s->q[0] = qnew;
        s->q[1] = nullptr;
        s->cur = 0;
        if (!s->inplace) {
            if (hipMalloc((void**) &s->q[1], sizeof(double) * s->pstride * (size_t) cap) != hipSuccess)
                return fail(HIPCLAW_ENOMEM, "hipclaw_regrid: device allocation failed%s");
            HIP_TRY(hipMemset(s->q[1], 0, sizeof(double) * s->pstride * (size_t) cap));
        }
    }
    // a saved state names old slots (and is sized for the old capacity)
    s->saved = false;
    s->save_lazy = false;
    if (cap > s->max_patches) {
        hipFree(s->qsave);
        s->qsave = nullptr;
        if (int rc = grow_patch_buffers(s, cap)) return rc;
    } else if (s->reg) {
        HIP_TRY(hipMemset(s->reg, 0, sizeof(double) * s->rstride * (size_t) s->max_patches));
    }
    // every program and the dt scale name old slots; the graph names old buffers
    hipFree(s->dtscale);
    s->dtscale = nullptr;
    for (GhostProgram& g : s->gp) {
        g.nops = 0;
        g.nlaunch[0] = g.nlaunch[1] = 0;
        g.layouts_built = false;
        g.set = false;
    }
    for (SyncProgram& y : s->sp) { y.ngroups = 0; y.nresets = 0; y.set = false; }
    s->lmax = s->lmin - 1;
    s->gp0_set = false;
    s->gp0_ops.clear();
    s->sp0_ops.clear();
    s->fused_ok = false;
    s->fsync = false;
    s->pull_ok = false;
    if (s->gexec) { hipGraphExecDestroy(s->gexec); s->gexec = nullptr; }
    if (s->graph) { hipGraphDestroy(s->graph); s->graph = nullptr; }
    s->graph_steps = 0;
    return hipclaw_set_geometry(s, npatch_new, dx, dy);
}

}  // extern "C"
