/*
 * hipclaw: C ABI of the MI355X batched patch-update backend.
 *
 * Plain pointers and sizes only.  Every entry point returns 0 on success or a
 * HIPCLAW_E* code; hipclaw_last_error() gives the message for the calling
 * thread.  Patch data keep the reference's FArrayBox layout: for each patch
 * q(1-mbc:mx+mbc, 1-mbc:my+mbc, meqn), i fastest, then j, then m
 * (src/patches/clawpatch/fclaw2d_clawpatch.hpp:42-80).
 *
 * The backend keeps every patch resident in HBM in one slab per solution
 * ("slots"); the host exchanges data with it only on request.  This replaces
 * the reference's per-batch host<->device copies
 * (src/solvers/fc2d_cudaclaw/cuda_source/cudaclaw_step2.cu:99-129, 245-266).
 */
#ifndef HIPCLAW_H
#define HIPCLAW_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    HIPCLAW_OK = 0,
    HIPCLAW_EINVAL = 1,   /* bad argument */
    HIPCLAW_ENOMEM = 2,   /* device allocation failed */
    HIPCLAW_EDEVICE = 3,  /* HIP runtime or kernel launch error */
    HIPCLAW_ENODEV = 4    /* no usable GPU */
};

/* Riemann solvers compiled into the backend (device functors, FP32 and FP64).
 * With those an application registers (hipclaw_solver.h, numbers from
 * HIPCLAW_RP_USER on) they replace the device function pointers cuda_rpn2 /
 * cuda_rpt2 of fc2d_cudaclaw_vtable_t
 * (src/solvers/fc2d_cudaclaw/fc2d_cudaclaw.h:159-162). */
enum {
    HIPCLAW_RP_ADVECTION = 0, /* clawpack46_rpn2adv.f / _rpt2adv.f, maux = 2 (u, v) */
    HIPCLAW_RP_SHALLOW = 1,   /* clawpack46_rpn2.f / _rpt2.f, Roe + entropy fix */
    HIPCLAW_RP_EULER = 2      /* clawpack46_rpn2_euler4.f / _rpt2_euler4.f, Roe + entropy fix */
};

enum {
    HIPCLAW_FP64 = 0,      /* all FP64; bit-faithful to the CPU Clawpack-4.6 path */
    HIPCLAW_FP32_BULK = 1  /* FP32 Riemann/limiter/transverse, FP64 q, fluxes, CFL, mass */
};

typedef struct hipclaw_config {
    int solver;      /* HIPCLAW_RP_* */
    int precision;   /* HIPCLAW_FP64 or HIPCLAW_FP32_BULK */
    int mx, my, mbc; /* cells per patch (mx, my >= 8, any size as the reference's
                        cudaclaw_flux2.cu:278-281; coarse/fine exchanges and time sync
                        need mx == my) and ghost layers (mbc = 2) */
    int meqn, maux, mwaves;
    int order[2];    /* Clawpack method(2:3): order and transverse propagation */
    int mthlim[10];  /* limiter per wave family (mwaves <= 10, FC2D_CUDACLAW_MWAVES) */
    int mthbc[4];    /* physical BC per domain face: 1 extrapolation, 3 wall */
    double grav;     /* shallow water */
    double gamma;    /* Euler */
    int tile;        /* cells per tile side (8, 16, 32; at most mx and my); 0 = tuned
                        default.  A tile that does not divide the patch overlaps its
                        last tile with the one before */
    int time_sync;   /* keep flux registers for the conservation fix-up
                        (fclaw_options time-sync, src/fclaw_options.c:110); needs mx == my */
    int subcycle;    /* keep the state before each update and time-interpolated
                        copies for local time stepping (fclaw_options subcycle) */
    double rp_params[8]; /* a registered solver's own constants (hipclaw_solver.h) */
} hipclaw_config_t;

/* Numbered ghost-fill and synchronisation programs per state; program 0 is
 * the one hipclaw_advance uses. */
#define HIPCLAW_MAX_PROGRAMS 32
/* Flag on a ghost op's slot: the slot's time-interpolated copy. */
#define HIPCLAW_TIME_INTERP_SLOT (1 << 30)

typedef struct hipclaw_state hipclaw_state_t;

/* Thread-local message for the last failing call. */
const char *hipclaw_last_error(void);

/* sizeof(hipclaw_config_t): lets a foreign-language binding check that its
 * copy of the struct matches the compiled library (no counterpart in the reference). */
size_t hipclaw_config_size(void);

/* Select the GPU and report its name.
 * Replaces fc2d_cudaclaw_initialize_GPUs
 * (src/solvers/fc2d_cudaclaw/cuda_source/cudaclaw_initialize.cu:37-74). */
int hipclaw_initialize_device(int device, char *name, size_t name_len);

/* Create a solver state with room for max_patches patches.  One allocation per
 * slab (q, its ping-pong partner, aux, geometry); no per-patch allocation.
 * Replaces fc2d_cudaclaw_allocate_buffers and the per-patch
 * cudaclaw_allocate_fluxes pool
 * (src/solvers/fc2d_cudaclaw/cuda_source/cudaclaw_allocate.cu:21-132) and
 * cudaclaw_set_method_parameters
 * (src/solvers/fc2d_cudaclaw/cuda_source/cudaclaw_flux2.cu:48-54). */
int hipclaw_create(const hipclaw_config_t *cfg, int max_patches, hipclaw_state_t **out);

/* Release all device memory of the state.
 * Replaces fc2d_cudaclaw_deallocate_buffers / cudaclaw_deallocate_fluxes. */
void hipclaw_destroy(hipclaw_state_t *s);

/* Tile size actually used (after defaults). */
int hipclaw_tile(const hipclaw_state_t *s);

/* Number of patches in use and their mesh widths (slot p has dx[p], dy[p]). */
int hipclaw_set_geometry(hipclaw_state_t *s, int npatch, const double *dx, const double *dy);

/* Host <-> device patch data for slots [first, first+count).
 * hipclaw_set_q / hipclaw_get_q replace the packing copies of
 * cudaclaw_step2_batch (cudaclaw_step2.cu:99-129, 245-266). */
int hipclaw_set_q(hipclaw_state_t *s, const double *q, int first, int count);
int hipclaw_get_q(hipclaw_state_t *s, double *q, int first, int count);
int hipclaw_set_aux(hipclaw_state_t *s, const double *aux, int first, int count);

/* Ghost-fill program: nops ops of 8 int32 each (type, dst, src, which, ti, tj,
 * bc, 0), grouped into 6 phases by phase_start[0..6]:
 *   0 average fine->coarse ghost, 1 copy same level, 2 physical BC on coarse
 *   levels, 3 interpolate coarse->fine ghost, 4 physical BC on fine levels,
 *   5 physical BC on all levels
 * (the order of fclaw2d_ghost_update_async, src/fclaw2d_ghost_fill.c:611-830).
 * Every op is checked on the host first: a bad one returns HIPCLAW_EINVAL and
 * leaves the installed program as it was.  The library keeps its own copy, so
 * the caller may reuse ops on return. */
int hipclaw_set_ghost_program(hipclaw_state_t *s, const int *ops, int nops,
                              const int *phase_start);

/* Run the ghost-fill program on the current solution.
 * Replaces the per-face host callbacks of fclaw2d_ghost_update
 * (src/fclaw2d_ghost_fill.c:838-856) and the Fortran kernels
 * fclaw2d_clawpatch46_fort_{copy,average,interpolate}_{face,corner}. */
int hipclaw_ghost_fill(hipclaw_state_t *s);

/* One explicit step of the listed patches (the batch of
 * fclaw2d_update_single_step), in place, at time step dt.  *maxcfl receives
 * the batch's maximum Courant number (FP64).
 * Replaces cudaclaw_step2_batch
 * (src/solvers/fc2d_cudaclaw/cuda_source/cudaclaw_step2.cu:49-269). */
int hipclaw_step_batch(hipclaw_state_t *s, const int *slots, int nslots, double t,
                       double dt, double *maxcfl);

/* Most slots one hipclaw_step_batch_noalloc call takes. */
#define HIPCLAW_NOALLOC_MAX_SLOTS 65536

/* The batched update of hipclaw_step_batch on caller-owned device memory,
 * allocating nothing: no hipMalloc, hipFree or allocating library call, so
 * callers that cannot allocate (code under graph capture, tight serving loops,
 * harnesses that forbid it) can drive it.  No counterpart in the reference.
 *   q_in, q_out : device, npatch slabs in the layout above; the interior cells
 *                 of the listed slots of q_out receive the updated state (their
 *                 ghost cells are not written).  q_in == q_out only when one
 *                 tile covers a patch (cfg->tile or its default).
 *   aux         : device, npatch aux slabs, exactly when cfg->maux > 0.
 *   dx, dy      : device, npatch doubles.
 *   slots       : host, nslots <= HIPCLAW_NOALLOC_MAX_SLOTS indices < npatch.
 *   stream      : the hipStream_t to run on (NULL: the default stream); the
 *                 call returns after the stream has finished its work.
 *   maxcfl      : the largest Courant number of the batch (may be NULL).
 * The slot list and the CFL word are staged in one fixed device scratch array
 * of the library (about 256 KiB per GPU, held until the process ends), cleared
 * by every call: calls that overlap on one GPU race on it.  No time-sync
 * registers and no per-slot dt scale.  Make the device current first (the
 * scratch address is looked up once per device, at hipclaw_initialize_device
 * or on the first call). */
int hipclaw_step_batch_noalloc(const hipclaw_config_t *cfg, int npatch, const double *q_in,
                               double *q_out, const double *aux, const double *dx,
                               const double *dy, const int *slots, int nslots, double dt,
                               void *stream, double *maxcfl);

/* Ghost-fill program number k (0 <= k < HIPCLAW_MAX_PROGRAMS); program 0 is
 * hipclaw_set_ghost_program's.  Programs for fclaw2d_ghost_update over a
 * partial level range with time interpolation (subcycling) reference the
 * coarser level's time-interpolated copies with HIPCLAW_TIME_INTERP_SLOT. */
int hipclaw_set_ghost_program_k(hipclaw_state_t *s, int k, const int *ops, int nops,
                                const int *phase_start);
int hipclaw_ghost_fill_k(hipclaw_state_t *s, int k);

/* Time interpolation of the listed patches for the ghost fill of their finer
 * neighbours at an intermediate time (config subcycle = 1): copies =
 * qlast + alpha (q - qlast) on the interior band of width mbc, where qlast is
 * the state saved by hipclaw_step_batch before the patch's last update.
 * Replaces fclaw2d_timeinterp / clawpatch_setup_timeinterp
 * (src/patches/clawpatch/fclaw2d_clawpatch.cpp:474-526). */
int hipclaw_timeinterp(hipclaw_state_t *s, const int *slots, int nslots, double alpha);

/* Time-synchronisation program number k, with the (slot, face) register
 * resets applied after its corrections (nresets pairs).  Registers add up
 * over substeps when the state was created with subcycle = 1; otherwise every
 * update overwrites them and program 0 needs no resets. */
int hipclaw_set_sync_program_k(hipclaw_state_t *s, int k, const int *ops, int nops,
                               const int *resets, int nresets);
int hipclaw_time_sync_k(hipclaw_state_t *s, int k);

/* Time-synchronisation program (config time_sync = 1): nops ops of 8 int32
 * (type, dst, src, face, ti, tj, 0, 0), type 0 = same-size correction of dst
 * from neighbour src across dst's face, 1 = fine-to-coarse correction of
 * coarse dst from fine src (the first fine child of coarse cell (ic, jc) is
 * (2ic-1+ti, 2jc-1+tj)).  Ops of one destination are contiguous, same-size
 * faces first, as correct_coarse_cells applies them
 * (src/fclaw2d_time_sync.c:118-158). */
int hipclaw_set_sync_program(hipclaw_state_t *s, const int *ops, int nops);

/* Apply the conservation fix-up from the flux registers of the last update of
 * every patch.  Replaces fclaw2d_time_sync (src/fclaw2d_time_sync.c:161-202)
 * with clawpack46_fort_time_sync_f2c / _samesize
 * (src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_time_sync.f:249-700). */
int hipclaw_time_sync(hipclaw_state_t *s);

/* Gradient refinement criterion of every patch (tag4refinement /
 * tag4coarsening with the gradient criterion,
 * src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_tag4refinement.f,
 * src/patches/clawpatch/fort_user/fclaw2d_clawpatch_gradient_exceeds_th.f90):
 * exceeds_refine[p] / exceeds_coarsen[p] = 1 if some interior cell of patch p
 * has |grad q(ivar)| above the refine / coarsen threshold (always 1 for a
 * negative refine threshold, "always refine", and for a coarsen threshold
 * <= 0, "never coarsen": src/patches/clawpatch/fclaw2d_clawpatch.cpp:813-817,
 * 890-911).  ivar is 0-based. */
int hipclaw_tag_patches(hipclaw_state_t *s, int ivar, double refine_threshold,
                        double coarsen_threshold, int *exceeds_refine, int *exceeds_coarsen);

/* Move the solution onto a new set of npatch_new patches (the repopulate step
 * of fclaw2d_regrid, src/fclaw2d_regrid.c:95-230): patch p is a copy of old
 * patch src[4p] (kind 0), interpolated from old parent src[4p] as child
 * src[4p+1] (kind 1, fclaw2d_clawpatch46_fort_interpolate2fine), or the
 * average of old children src[4p..4p+3] (kind 2,
 * fclaw2d_clawpatch46_fort_average2coarse) -- the provenance of
 * fclaw2d_hip_forest_adapt.  New mesh widths dx, dy.  Buffers grow as needed;
 * ghost and sync programs, a captured graph and aux are invalidated (set them
 * again), registers start at zero. */
int hipclaw_regrid(hipclaw_state_t *s, int npatch_new, const int *kind, const int *src,
                   const double *dx, const double *dy);

/* Ghost-patch messages: what one rank sends another for each patch that
 * borders it, straight from and into the device slabs.  Replaces
 * clawpatch_local_ghost_pack / clawpatch_remote_ghost_unpack
 * (src/patches/clawpatch/fclaw2d_clawpatch.cpp:1019-1215) with
 * fclaw2d_clawpatch46_fort_local_ghost_pack
 * (src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_local_ghost_pack.f) and,
 * with time_sync, fclaw2d_clawpatch_time_sync_pack_registers
 * (src/patches/clawpatch/fclaw2d_clawpatch_conservation.c:112-195), byte for
 * byte the reference's message layout (refratio 2, no manifold area, no
 * extra fields): *elems = doubles per patch (clawpatch_ghost_pack_elems,
 * src/patches/clawpatch/fclaw2d_clawpatch.cpp:1020-1060).  buf is DEVICE
 * memory holding n consecutive messages (checked against the allocation that
 * contains it); slot k's message is
 * buf[k * elems ...].  time_interp = 1 packs from / unpacks into the
 * time-interpolated copies (config subcycle = 1).  Unpacking writes q and
 * the registers; edge lengths and areas follow from dx, dy. */
int hipclaw_ghost_pack_elems(const hipclaw_state_t *s, long long *elems);
int hipclaw_ghost_pack(hipclaw_state_t *s, const int *slots, int n, int time_interp,
                       double *buf);
int hipclaw_ghost_unpack(hipclaw_state_t *s, const int *slots, int n, int time_interp,
                         const double *buf);

/* Per-slot time step: every later update of slot p advances it by
 * dt * scale[p] (the dt of hipclaw_advance / hipclaw_step_batch), so the
 * patches of an ensemble of independent runs -- each with its own stable dt,
 * in one forest of isolated blocks (fclaw2d_hip_forest_set_isolated) -- share
 * one batched launch; the CFL number is then the maximum over all runs.
 * scale = NULL restores a common dt.  A regrid drops the scale (slots
 * change).  No counterpart in the reference, which runs one dt per level. */
int hipclaw_set_dt_scale(hipclaw_state_t *s, const double *scale);

/* Copy the current solution of every patch (ghost cells included) to the
 * saved state on the device, or the saved state back: clawpatch_save_step /
 * clawpatch_restore_step (src/patches/clawpatch/fclaw2d_clawpatch.cpp:446-470)
 * for all patches at once, as save_time_step / restore_time_step of the run
 * loop use them to reject a step whose Courant number exceeds max_cfl
 * (src/fclaw2d_run.c:47-80).  Restoring without a saved state, or after a
 * regrid (the slots changed), is an error.  With two slabs (tile smaller
 * than the patch) the saved state stays in its slab without a copy while
 * only a one-step advance runs (it writes the other slab); any other call
 * that may write that slab copies it first.  A restore may so make the
 * other slab current: a graph captured before then refuses to replay. */
int hipclaw_save_step(hipclaw_state_t *s);
int hipclaw_restore_step(hipclaw_state_t *s);

/* Launch layout of the ghost fill, the same ops either way: 0 = one launch
 * per phase (average + copy, physical BC, interpolate, physical BC), 1 = two
 * launches of per-destination groups (a patch's exchange ops, then its
 * physical BC), 2 = fused: one launch per step of per-patch groups in which
 * interpolation recomputes the coarse patch's face ghost cells it reads from
 * their sources (same-size neighbour, fine children, physical BC), and, with
 * time sync, the corrected edge values go to a side buffer that the next
 * update stages (the last step's are written into q at the end of
 * hipclaw_advance), so a step is two launches.  The default is 2; where the
 * program cannot be fused (time-interpolated slots, subcycling)
 * hipclaw_ghost_layout reports and uses 1 for mx >= 32, else 0.  No
 * counterpart in the reference, which fills ghost cells patch by patch on
 * the host (src/fclaw2d_ghost_fill.c:611-830) and then synchronises
 * (src/fclaw2d_time_sync.c:157-202). */
int hipclaw_set_ghost_layout(hipclaw_state_t *s, int layout);
int hipclaw_ghost_layout(const hipclaw_state_t *s);

/* Waves per workgroup of the 16 x 16 update tile: 2, 3 or 4 (0 = the
 * default: 2 for the FP32-bulk form, 4 for FP64).  A sweep has six row
 * groups: two waves take three each, three take two, four take two, two, one
 * and one; fewer waves per workgroup hold fewer staging registers, and more
 * workgroups share a CU when the tile's LDS allows.
 * hipclaw_waves returns the waves the next update uses (0 for tiles without
 * a choice).  No counterpart in the reference (128 threads per patch). */
int hipclaw_set_waves(hipclaw_state_t *s, int waves);
int hipclaw_waves(const hipclaw_state_t *s);

/* Most trials of one hipclaw_autotune call. */
#define HIPCLAW_AUTOTUNE_MAX_TRIALS 16

/* Autotuner: time nsteps full steps (update, ghost fill, time sync) of dt on
 * the current patches for every tile size the state's solver, precision and
 * patch size allow, with 2, 3 and 4 waves per workgroup for 16 x 16 tiles, in
 * every ghost-fill layout the program allows, each trial from the state saved
 * at entry and restored after it (the saved state of hipclaw_save_step is
 * overwritten), and keep the fastest.  tiles[k], waves[k], layouts[k], ms[k]
 * (room for HIPCLAW_AUTOTUNE_MAX_TRIALS)
 * receive each trial and its mean step time, *ntrials their count.  A
 * subcycling state (levels set) is timed on subcycled coarse steps.  No
 * counterpart in the reference, whose CUDA kernels have a fixed block size. */
int hipclaw_autotune(hipclaw_state_t *s, double dt, int nsteps, int *tiles, int *waves,
                     int *layouts, double *ms, int *ntrials);

/* Cell values of slots first..first+count-1 as float, in the order of the
 * VTK "meqn" cell array (per patch: j = 1..my, i = 1..mx, then the meqn
 * fields), converted on the device so that half the bytes cross PCIe.
 * Replaces fclaw2d_output_vtk_value_cb
 * (src/patches/clawpatch/fclaw2d_clawpatch_output_vtk.c:781-808). */
int hipclaw_vtk_values(hipclaw_state_t *s, int first, int count, float *out);

/* Copy the time-sync registers of slots first..first+count-1 to host memory,
 * rstride = 16 mx meqn doubles per slot in the layout
 * reg[((face * 4 + kind) * meqn + m) * mx + pos] (kind 0..3 = fm, fp, edge
 * flux at the first interior cell, edge flux at the first ghost cell).  The
 * counterpart of reading fclaw2d_clawpatch_get_registers
 * (src/patches/clawpatch/fclaw2d_clawpatch.cpp) on the host. */
int hipclaw_get_registers(hipclaw_state_t *s, int first, int count, double *out);

/* nsteps of (update every patch at every level, ghost fill, and with
 * time_sync the conservation fix-up), all resident on the GPU; *maxcfl
 * receives the maximum over the steps.  Equivalent to
 * fclaw2d_advance_all_levels with subcycle = F
 * (src/fclaw2d_advance.c:247-357). */
int hipclaw_advance(hipclaw_state_t *s, double dt, int nsteps, double *maxcfl);

/* Same as hipclaw_advance but does not wait for the GPU and does not return
 * the CFL number; hipclaw_sync() waits.  For timing and graph replay. */
int hipclaw_advance_async(hipclaw_state_t *s, double dt, int nsteps);
int hipclaw_sync(hipclaw_state_t *s);

/* Capture nsteps of hipclaw_advance_async in a HIP graph (replaces any graph
 * captured before) and replay it.  The graph starts from the slab current at
 * its capture; a replay when the other slab is current (after an odd number
 * of steps, or a restore) is refused with HIPCLAW_EINVAL.  Installing a ghost
 * or sync program, the levels, or geometry for another number of patches
 * drops the graph (its launches hold the old tables and counts). */
int hipclaw_capture_graph(hipclaw_state_t *s, double dt, int nsteps);
int hipclaw_replay_graph(hipclaw_state_t *s);

/* Compensated FP64 sums of q*dx*dy over interior cells, per field:
 * sum[m] = hi + lo of a double-double accumulation.
 * Replaces fclaw2d_clawpatch46_fort_conservation_check
 * (src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_diagnostics.f:13-55). */
int hipclaw_mass(hipclaw_state_t *s, double *sum_hi, double *sum_lo);

/* Timing of the last hipclaw_advance_async/replay region, measured with HIP
 * events on the state's stream: total milliseconds and milliseconds spent in
 * the patch-update kernel alone (per launch average, from events bound to each
 * update kernel's own start and end, as a kernel trace reports it). */
int hipclaw_timing(hipclaw_state_t *s, int nsteps, double *total_ms, double *step_kernel_ms);

/* Levels of slots 0..npatch-1 (npatch = the current patch count), for the
 * subcycled advance.  Regrid and a change of the patch count forget them. */
int hipclaw_set_levels(hipclaw_state_t *s, const int *level, int npatch);

/* ncoarse calls of fclaw2d_advance_all_levels(dt) with subcycle = T
 * (src/fclaw2d_advance.c:54-246, the level loop of :112-170 and the ghost
 * updates / time syncs of :292-330): level L takes 2^(L - lmin) steps of
 * dt / 2^(L - lmin), the state before each step is kept, coarse levels are
 * time-interpolated for the finer levels' ghost cells, and partial ghost
 * fills (program 1 + ti - lmin: levels ti+1..lmax with ti time-interpolated)
 * and time syncs (program L - lmin: levels L..lmax) run between finest steps,
 * all enqueued on the device with no host synchronisation; *maxcfl receives
 * the largest Courant number of every substep.  Needs a state created with
 * subcycle = 1, hipclaw_set_levels, and ghost (and, with time sync, sync)
 * programs 0 .. lmax - lmin. */
int hipclaw_advance_subcycle(hipclaw_state_t *s, double dt, int ncoarse, double *maxcfl);
int hipclaw_advance_subcycle_async(hipclaw_state_t *s, double dt, int ncoarse);

/* Capture ncoarse subcycled coarse steps into a graph (replayed with
 * hipclaw_replay_graph). */
int hipclaw_capture_subcycle_graph(hipclaw_state_t *s, double dt, int ncoarse);

#ifdef __cplusplus
}
#endif
#endif
