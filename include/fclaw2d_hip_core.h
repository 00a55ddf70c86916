/*
 * The parts of ForestClaw's core C interface that the batched patch update
 * plugs into, declared for this library (the reference's own headers need
 * p4est/libsc).  Names, argument order and meaning follow the reference:
 *
 *   fclaw2d_patch_t / fclaw2d_block_t / fclaw2d_domain_t
 *                      src/forestclaw2d.h:97-170 (fields this path uses)
 *   fclaw2d_patch_vtable_t, single_step_update typedef
 *                      src/fclaw2d_patch.h:937-943, 1413-1570
 *   fclaw2d_single_step_buffer_data_t, fclaw2d_update_single_step
 *                      src/fclaw2d_update_single_step.h:47-95
 *   fclaw2d_ghost_update
 *                      src/fclaw2d_ghost_fill.h (src/fclaw2d_ghost_fill.c:838-856)
 *   fclaw2d_clawpatch_{grid,soln,aux}_data
 *                      src/patches/clawpatch/fclaw2d_clawpatch.h
 *
 * Patch data live on the GPU; the clawpatch accessors return a host view of a
 * patch that fc2d_hipclaw_download / fc2d_hipclaw_upload keep in step with it.
 */
#ifndef FCLAW2D_HIP_CORE_H
#define FCLAW2D_HIP_CORE_H

#include <stddef.h>

#include "fclaw2d_hip_forest.h"

#ifdef __cplusplus
extern "C" {
#endif

struct fclaw2d_global;
struct fc2d_hipclaw_vtable;
struct fc2d_hipclaw_options;
struct hipclaw_state;

typedef struct fclaw2d_patch {
    int level;          /* refinement level */
    double xlower, xupper, ylower, yupper;  /* in [0,1] block coordinates */
    int blockno;        /* owning block */
    int patchno;        /* index within the block */
    int slot;           /* index of the patch's data on the device */
    void *user;
} fclaw2d_patch_t;

typedef struct fclaw2d_block {
    int num_patches;
    int patchno_offset;        /* slot of the block's first patch */
    fclaw2d_patch_t *patches;
} fclaw2d_block_t;

typedef struct fclaw2d_domain {
    int num_blocks;
    fclaw2d_block_t *blocks;
    int local_num_patches;
    int global_minlevel, global_maxlevel;
    int nbx, nby, periodic_x, periodic_y;
    double ax, bx, ay, by;     /* physical extent of the brick */
    fclaw2d_hip_forest_t *forest;
} fclaw2d_domain_t;

typedef double (*fclaw2d_patch_single_step_update_t)(struct fclaw2d_global *glob,
                                                     fclaw2d_patch_t *this_patch,
                                                     int this_block_idx, int this_patch_idx,
                                                     double t, double dt, void *user);

typedef void (*fclaw2d_patch_initialize_t)(struct fclaw2d_global *glob,
                                           fclaw2d_patch_t *this_patch,
                                           int this_block_idx, int this_patch_idx);

typedef void (*fclaw2d_patch_physical_bc_t)(struct fclaw2d_global *glob,
                                            fclaw2d_patch_t *this_patch,
                                            int this_block_idx, int this_patch_idx,
                                            double t, double dt, int intersects_bc[],
                                            int time_interp);

typedef struct fclaw2d_patch_vtable {
    fclaw2d_patch_initialize_t initialize;  /* qinit */
    fclaw2d_patch_initialize_t setup;       /* setaux */
    fclaw2d_patch_physical_bc_t physical_bc;
    fclaw2d_patch_single_step_update_t single_step_update;
    int is_set;
} fclaw2d_patch_vtable_t;

typedef struct fclaw2d_clawpatch_options {
    int mx, my, mbc, meqn, maux;
} fclaw2d_clawpatch_options_t;

/* refinement criteria codes (src/patches/clawpatch/fclaw2d_clawpatch_options.h:48-56) */
#define FCLAW_REFINE_CRITERIA_VALUE 0
#define FCLAW_REFINE_CRITERIA_DIFFERENCE 1
#define FCLAW_REFINE_CRITERIA_MINMAX 2
#define FCLAW_REFINE_CRITERIA_GRADIENT 3
#define FCLAW_REFINE_CRITERIA_USER 4

typedef struct fclaw2d_global {
    fclaw2d_domain_t *domain;
    int count_single_step;         /* patch updates so far (src/fclaw_timer.c:184) */
    int count_ghost_update;
    double curr_time, curr_dt;
    fclaw2d_patch_vtable_t patch_vt;
    fclaw2d_clawpatch_options_t clawpatch_opt;
    struct fc2d_hipclaw_vtable *hipclaw_vt;
    struct fc2d_hipclaw_options *hipclaw_opt;
    struct hipclaw_state *state;   /* device-resident solution */
    double *qhost, *auxhost;       /* host views [slot][field][j][i] */
    int time_sync;                 /* fclaw_options time-sync (src/fclaw_options.c:110) */
    int subcycle;                  /* fclaw_options subcycle (src/fclaw_options.c:99) */
    int advance_one_step;          /* fclaw_options advance-one-step (src/fclaw_options.c:91) */
    /* regrid options (src/fclaw_options.c:190-221) and the clawpatch
     * threshold-variable, 1..meqn (src/patches/clawpatch/fclaw2d_clawpatch_options.c:112) */
    int minlevel, maxlevel;
    double refine_threshold, coarsen_threshold;
    int threshold_variable;
    /* clawpatch refinement-criteria, FCLAW_REFINE_CRITERIA_* below (default
     * minmax, src/patches/clawpatch/fclaw2d_clawpatch_options.c:107-110);
     * fclaw2d_initialize and fclaw2d_regrid tag with the gradient criterion
     * only and stop with a message for any other value */
    int refinement_criteria;
    int count_amr_regrid, count_amr_new_domain;   /* src/fclaw2d_regrid.c:321, 333 */
    const char *prefix;            /* output file prefix, NULL = "fort" (src/fclaw_options.c:177) */
    /* time stepping of fclaw2d_run (src/fclaw_options.c:62-98, 196-198); the
     * defaults of fclaw2d_global_new are the reference's */
    double initial_dt, max_cfl, desired_cfl, tfinal;
    int reduce_cfl, use_fixed_dt, outstyle, nout, nstep, outstyle_uses_maxlevel;
    int regrid_interval;
    int output;                    /* fclaw_options output: write frames (src/fclaw_options.c:122) */
    int count_retake;              /* steps rejected for max_cfl and retaken */
    /* fclaw_options conservation-check (src/fclaw_options.c:233): the last
     * gathered total of each field and the total at t = 0 */
    int conservation_check;
    double cons_mass[4], cons_mass0[4];
} fclaw2d_global_t;

/** Buffer contract between fclaw2d_update_single_step and single_step_update
 *  (src/fclaw2d_update_single_step.h:47-55). */
typedef struct fclaw2d_single_step_buffer_data {
    int total_count;   /* patches updated by this call */
    int iter;          /* index of the current patch, incremented by the caller */
    void *user;        /* owned by the callee: allocated at iter == 0, freed at the last */
} fclaw2d_single_step_buffer_data_t;

/* sizeof of fclaw2d_global_t, fclaw2d_domain_t, fclaw2d_block_t, fclaw2d_patch_t
 * (out[0..3]) and of fc2d_hipclaw_options_t, fc2d_hipclaw_vtable_t (out[4..5]),
 * for bindings to check their struct copies against the compiled library. */
void fclaw2d_hip_struct_sizes(size_t out[6]);

/* ---- globals, domains ---- */
fclaw2d_global_t *fclaw2d_global_new(void);
void fclaw2d_global_destroy(fclaw2d_global_t *glob);

/* Domain over the brick [ax,bx] x [ay,by] of nbx x nby blocks whose leaves are
 * those of `forest` (the domain takes ownership of the forest). */
fclaw2d_domain_t *fclaw2d_domain_new_from_forest(fclaw2d_hip_forest_t *forest, int nbx, int nby,
                                                 int periodic_x, int periodic_y, double ax,
                                                 double bx, double ay, double by);
void fclaw2d_domain_destroy(fclaw2d_domain_t *domain);
void fclaw2d_global_store_domain(fclaw2d_global_t *glob, fclaw2d_domain_t *domain);

fclaw2d_patch_vtable_t *fclaw2d_patch_vt(fclaw2d_global_t *glob);

/* ---- the hot path's callers ---- */

/** Advance every patch at `level` by one explicit step; returns the maximum
 *  Courant number (src/fclaw2d_update_single_step.c:73-103). */
double fclaw2d_update_single_step(fclaw2d_global_t *glob, int level, double t, double dt);

/** Fill ghost cells of all patches at levels minlevel..maxlevel
 *  (src/fclaw2d_ghost_fill.c:838-856).  The level ranges are those the
 *  advance asks for: every level, or (time_interp, with subcycling) the
 *  levels above a time-interpolated one. */
void fclaw2d_ghost_update(fclaw2d_global_t *glob, int minlevel, int maxlevel,
                          double sync_time, int time_interp, int running);

/** Advance every level from t_curr to t_curr + dt, ending with a ghost
 *  update (and time sync) of every level (fclaw2d_advance_all_levels,
 *  src/fclaw2d_advance.c:247-357, with glob's subcycle / advance_one_step /
 *  time_sync).  Returns the maximum Courant number. */
double fclaw2d_advance_all_levels(fclaw2d_global_t *glob, double t_curr, double dt);

/* Time-interpolate the patches of `level` (fclaw2d_timeinterp,
 * src/fclaw2d_timeinterp.c); needs glob->subcycle. */
void fclaw2d_timeinterp(fclaw2d_global_t *glob, int level, double alpha);

/* ---- mesh adaptation ----
 * fclaw2d_initialize (src/fclaw2d_initialize.c:75-240): device state for the
 * current domain, setaux + qinit on every patch, ghost update, then for each
 * level minlevel..maxlevel-1 tag (no coarsening), refine and initialize the
 * new domain again, until no patch refines.  Ghost cells are filled before
 * each tagging.
 * fclaw2d_regrid (src/fclaw2d_regrid.c:235-340): gradient tags on the device
 * (refine_threshold / coarsen_threshold on threshold_variable), forest adapt
 * (refine below maxlevel, coarsen families above minlevel, 2:1 balance), the
 * solution moved to the new leaves on the device (copy, interpolate2fine or
 * average2coarse), setaux on the host, then a ghost update of every level.
 * Nothing moves when the leaves do not change. */
void fclaw2d_initialize(fclaw2d_global_t *glob);
void fclaw2d_regrid(fclaw2d_global_t *glob);

/* ---- the run loop ----
 * fclaw2d_run (src/fclaw2d_run.c:516-542) with output styles 0, 1, 3 and 4
 * (src/fclaw2d_run.c:86-300, 306-514): steps of fclaw2d_advance_all_levels
 * with dt set from the previous step's Courant number (dt * desired_cfl /
 * maxcfl), a step whose Courant number exceeds max_cfl rejected and retaken
 * from the saved state (unless use_fixed_dt), a regrid every
 * regrid_interval steps, and a frame (fort.qXXXX and/or
 * <prefix>_frame_XXXX.vtu, as the solver options ascii_out / vtk_out ask)
 * at every output time when glob->output is set.  Other styles exit with
 * the reference's message.  save / restore_time_step keep and put back the
 * state of every patch on the device (src/fclaw2d_run.c:47-80). */
void fclaw2d_run(fclaw2d_global_t *glob);

/* fclaw2d_diagnostics_gather (src/fclaw2d_diagnostics.c:122-190) with the
 * clawpatch conservation check
 * (src/patches/clawpatch/fclaw2d_clawpatch_diagnostics.c:159-228): when
 * glob->conservation_check is set, the compensated device sum of q dx dy of
 * every field (hipclaw_mass) goes to glob->cons_mass (and cons_mass0 when
 * init_flag), and "sum[m] = total |total - total at t=0|" is printed as the
 * reference prints it.  The error norms against an exact solution are not
 * provided. */
void fclaw2d_diagnostics_gather(fclaw2d_global_t *glob, int init_flag);
void fclaw2d_save_time_step(fclaw2d_global_t *glob);
void fclaw2d_restore_time_step(fclaw2d_global_t *glob);

/* ---- output ----
 * ASCII frames as the reference writes them: fort.tXXXX (time header) and
 * fort.qXXXX (every patch), fclaw2d_clawpatch_output_ascii
 * (src/patches/clawpatch/fclaw2d_clawpatch_output_ascii.c:112-160) with the
 * formats of src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_output_ascii.f.
 * The first downloads the device solution; the other two are its pieces
 * (time header, then one patch appended), return 0 or -1 on a file error. */
void fclaw2d_clawpatch_output_ascii(fclaw2d_global_t *glob, int iframe);
int fclaw2d_clawpatch_time_header_ascii_files(const char *fname_q, const char *fname_t,
                                              double time, int meqn, int maux, int ngrids);
int fclaw2d_clawpatch_output_ascii_patch(const char *fname_q, int mx, int my, int meqn, int mbc,
                                         double xlower, double ylower, double dx, double dy,
                                         const double *q, int patch_num, int level, int blockno,
                                         int mpirank);

/* VTK frames as the reference writes them: <prefix>_frame_XXXX.vtu, an
 * UnstructuredGrid with raw appended data (Position, connectivity, offsets,
 * types, mpirank, blockno, patchno, meqn), fclaw2d_vtk_write_file and
 * fclaw2d_clawpatch_output_vtk
 * (src/patches/clawpatch/fclaw2d_clawpatch_output_vtk.c:597-900).  The
 * callbacks fill one patch's points (3 doubles each, (mx+1)(my+1) of them)
 * or cell values (mx my meqn floats) into a.  fclaw2d_clawpatch_output_vtk
 * converts the values on the device (hipclaw_vtk_values); the callbacks below
 * read the host views.  Returns 0 or -1 on a file error. */
typedef void (*fclaw2d_vtk_patch_data_t)(fclaw2d_global_t *glob, fclaw2d_patch_t *this_patch,
                                         int this_block_idx, int this_patch_idx, char *a);
int fclaw2d_vtk_write_file(fclaw2d_global_t *glob, const char *basename, int mx, int my,
                           int meqn, double vtkspace, int vtkwrite,
                           fclaw2d_vtk_patch_data_t coordinate_cb,
                           fclaw2d_vtk_patch_data_t value_cb);
void fclaw2d_output_vtk_coordinate_cb(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                                      int patchno, char *a);
void fclaw2d_output_vtk_value_cb(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                                 int patchno, char *a);
void fclaw2d_clawpatch_output_vtk(fclaw2d_global_t *glob, int iframe);

/* Conservation fix-up after all levels reached the same time
 * (src/fclaw2d_time_sync.c:161-202); needs glob->time_sync set before
 * fc2d_hipclaw_allocate_buffers. */
void fclaw2d_time_sync(fclaw2d_global_t *glob, int minlevel, int maxlevel);

/* ---- clawpatch accessors (host views) ---- */
void fclaw2d_clawpatch_grid_data(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int *mx,
                                 int *my, int *mbc, double *xlower, double *ylower, double *dx,
                                 double *dy);
void fclaw2d_clawpatch_soln_data(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, double **q,
                                 int *meqn);
void fclaw2d_clawpatch_aux_data(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, double **aux,
                                int *maux);

#ifdef __cplusplus
}
#endif
#endif
