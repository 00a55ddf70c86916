/*
 * Host side of the batched patch update, in C: the reference's plugin surface
 * (declared in include/fclaw2d_hip_core.h and include/fc2d_hipclaw.h) on top
 * of the device C ABI (include/hipclaw.h).
 *
 *   fclaw2d_update_single_step : src/fclaw2d_update_single_step.c:32-103
 *   hipclaw_update (the single_step_update slot) : fc2d_cudaclaw.cpp:272-368
 *   fc2d_hipclaw_solver_initialize : fc2d_cudaclaw.cpp:405-465
 *   options defaults and checks : fc2d_cudaclaw_options.c:36-115
 *
 * Error behaviour follows the reference: a failing device call prints the
 * backend's message and exits with the error code
 * (src/solvers/fc2d_cudaclaw/fc2d_cudaclaw_check.h:3-11); option checks
 * return an error code.
 */
#include "../../include/fc2d_hipclaw.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void device_check(int rc, const char *what)
{
    if (rc != HIPCLAW_OK) {
        fprintf(stderr, "fc2d_hipclaw: %s failed: %s\n", what, hipclaw_last_error());
        exit(rc);
    }
}

/* ------------------------------------------------------------------ globals */

void fclaw2d_hip_struct_sizes(size_t out[6])
{
    out[0] = sizeof(fclaw2d_global_t);
    out[1] = sizeof(fclaw2d_domain_t);
    out[2] = sizeof(fclaw2d_block_t);
    out[3] = sizeof(fclaw2d_patch_t);
    out[4] = sizeof(fc2d_hipclaw_options_t);
    out[5] = sizeof(fc2d_hipclaw_vtable_t);
}

fclaw2d_global_t *fclaw2d_global_new(void)
{
    fclaw2d_global_t *g = (fclaw2d_global_t *) calloc(1, sizeof(fclaw2d_global_t));
    if (!g) return NULL;
    /* time stepping defaults of src/fclaw_options.c:62-98, 196-198 */
    g->initial_dt = 0.1;
    g->max_cfl = 1.0;
    g->desired_cfl = 0.9;
    g->reduce_cfl = 1;
    g->use_fixed_dt = 0;
    g->outstyle = 1;
    g->tfinal = 1.0;
    g->nout = 10;
    g->nstep = 1;
    g->regrid_interval = 1;
    g->threshold_variable = 1;
    g->refinement_criteria = FCLAW_REFINE_CRITERIA_MINMAX;
    return g;
}

void fclaw2d_global_destroy(fclaw2d_global_t *glob)
{
    if (!glob) return;
    fc2d_hipclaw_deallocate_buffers(glob);
    fclaw2d_domain_destroy(glob->domain);
    free(glob->hipclaw_vt);
    free(glob->hipclaw_opt);
    free(glob);
}

fclaw2d_domain_t *fclaw2d_domain_new_from_forest(fclaw2d_hip_forest_t *forest, int nbx, int nby,
                                                 int periodic_x, int periodic_y, double ax,
                                                 double bx, double ay, double by)
{
    fclaw2d_domain_t *d;
    int n, p, b, *lv, *gx, *gy;
    if (!forest) return NULL;
    n = fclaw2d_hip_forest_num_patches(forest);
    d = (fclaw2d_domain_t *) calloc(1, sizeof(*d));
    lv = (int *) malloc(sizeof(int) * n);
    gx = (int *) malloc(sizeof(int) * n);
    gy = (int *) malloc(sizeof(int) * n);
    if (!d || !lv || !gx || !gy) { free(d); free(lv); free(gx); free(gy); return NULL; }
    fclaw2d_hip_forest_leaves(forest, lv, gx, gy);
    d->num_blocks = nbx * nby;
    d->blocks = (fclaw2d_block_t *) calloc((size_t) d->num_blocks, sizeof(fclaw2d_block_t));
    d->local_num_patches = n;
    d->global_minlevel = fclaw2d_hip_forest_min_level(forest);
    d->global_maxlevel = fclaw2d_hip_forest_max_level(forest);
    d->nbx = nbx; d->nby = nby; d->periodic_x = periodic_x; d->periodic_y = periodic_y;
    d->ax = ax; d->bx = bx; d->ay = ay; d->by = by;
    d->forest = forest;
    /* leaves are ordered by block, then Morton index: count, then fill */
    for (p = 0; p < n; p++) {
        const int blk = (gx[p] >> lv[p]) + nbx * (gy[p] >> lv[p]);
        d->blocks[blk].num_patches++;
    }
    for (b = 0, p = 0; b < d->num_blocks; b++) {
        d->blocks[b].patchno_offset = p;
        d->blocks[b].patches = (fclaw2d_patch_t *) calloc((size_t) d->blocks[b].num_patches + 1,
                                                          sizeof(fclaw2d_patch_t));
        p += d->blocks[b].num_patches;
    }
    for (p = 0; p < n; p++) {
        const int L = lv[p], bxi = gx[p] >> L, byi = gy[p] >> L, blk = bxi + nbx * byi;
        const int k = p - d->blocks[blk].patchno_offset;
        fclaw2d_patch_t *pt = &d->blocks[blk].patches[k];
        const double w = ldexp(1.0, -L);
        pt->level = L;
        pt->xlower = (gx[p] - (bxi << L)) * w; pt->xupper = pt->xlower + w;
        pt->ylower = (gy[p] - (byi << L)) * w; pt->yupper = pt->ylower + w;
        pt->blockno = blk;
        pt->patchno = k;
        pt->slot = p;
    }
    free(lv); free(gx); free(gy);
    return d;
}

void fclaw2d_domain_destroy(fclaw2d_domain_t *d)
{
    int b;
    if (!d) return;
    for (b = 0; b < d->num_blocks; b++) free(d->blocks[b].patches);
    free(d->blocks);
    fclaw2d_hip_forest_destroy(d->forest);
    free(d);
}

void fclaw2d_global_store_domain(fclaw2d_global_t *glob, fclaw2d_domain_t *domain)
{
    glob->domain = domain;
}

fclaw2d_patch_vtable_t *fclaw2d_patch_vt(fclaw2d_global_t *glob) { return &glob->patch_vt; }

/* ------------------------------------------------------------------ clawpatch views */

static size_t patch_size(const fclaw2d_global_t *glob, int fields)
{
    const fclaw2d_clawpatch_options_t *o = &glob->clawpatch_opt;
    return (size_t) fields * (o->mx + 2 * o->mbc) * (o->my + 2 * o->mbc);
}

void fclaw2d_clawpatch_grid_data(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int *mx,
                                 int *my, int *mbc, double *xlower, double *ylower, double *dx,
                                 double *dy)
{
    const fclaw2d_domain_t *d = glob->domain;
    const fclaw2d_clawpatch_options_t *o = &glob->clawpatch_opt;
    const double wx = (d->bx - d->ax) / d->nbx, wy = (d->by - d->ay) / d->nby;
    const int bxi = patch->blockno % d->nbx, byi = patch->blockno / d->nbx;
    *mx = o->mx; *my = o->my; *mbc = o->mbc;
    *xlower = d->ax + (bxi + patch->xlower) * wx;
    *ylower = d->ay + (byi + patch->ylower) * wy;
    *dx = (patch->xupper - patch->xlower) * wx / o->mx;
    *dy = (patch->yupper - patch->ylower) * wy / o->my;
}

void fclaw2d_clawpatch_soln_data(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, double **q,
                                 int *meqn)
{
    *meqn = glob->clawpatch_opt.meqn;
    *q = glob->qhost ? glob->qhost + patch_size(glob, *meqn) * patch->slot : NULL;
}

void fclaw2d_clawpatch_aux_data(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, double **aux,
                                int *maux)
{
    *maux = glob->clawpatch_opt.maux;
    *aux = (glob->auxhost && *maux > 0) ? glob->auxhost + patch_size(glob, *maux) * patch->slot
                                        : NULL;
}

/* ------------------------------------------------------------------ options */

void fc2d_hipclaw_options_default(fc2d_hipclaw_options_t *opt, int mwaves)
{
    int k;
    memset(opt, 0, sizeof(*opt));
    opt->mwaves = mwaves;
    opt->order[0] = 2; opt->order[1] = 2;
    for (k = 0; k < FC2D_HIPCLAW_MWAVES; k++) opt->mthlim[k] = 1;
    for (k = 0; k < 4; k++) opt->mthbc[k] = 1;
    opt->buffer_len = 1;
    opt->precision = HIPCLAW_FP64;
    opt->grav = 1.0;
    opt->gamma = 1.4;
    opt->is_registered = 1;
}

int hipclaw_check_parameters(int mwaves)
{
    if (mwaves > FC2D_HIPCLAW_MWAVES) {
        fprintf(stderr, "fc2d_hipclaw : mwaves > FC2D_HIPCLAW_MWAVES (%d)\n", FC2D_HIPCLAW_MWAVES);
        return 1;
    }
    return 0;
}

int fc2d_hipclaw_check(const fc2d_hipclaw_options_t *opt)
{
    if (hipclaw_check_parameters(opt->mwaves)) return 1;
    if (opt->buffer_len < 1) return 1;
    if (opt->mcapa != 0 || opt->use_fwaves != 0) {
        fprintf(stderr, "fc2d_hipclaw : mcapa and f-waves are not supported\n");
        return 1;
    }
    /* mthbc 1 (extrapolation) and 3 (wall) are applied on physical faces; 2
     * (periodic) is left to the domain's topology.  The reference stops on 0
     * and on 2 at a physical face (clawpack46_bc2_default.f:22-25, 41-43); the
     * ghost program refuses 2 there. */
    {
        int k;
        for (k = 0; k < 4; k++)
            if (opt->mthbc[k] != 1 && opt->mthbc[k] != 2 && opt->mthbc[k] != 3) {
                fprintf(stderr, "fc2d_hipclaw : mthbc(%d) = %d is not 1, 2 or 3\n", k + 1,
                        opt->mthbc[k]);
                return 1;
            }
    }
    return 0;
}

void fc2d_hipclaw_options_store(fclaw2d_global_t *glob, fc2d_hipclaw_options_t *opt)
{
    if (!glob->hipclaw_opt) glob->hipclaw_opt = (fc2d_hipclaw_options_t *) malloc(sizeof(*opt));
    *glob->hipclaw_opt = *opt;
}

fc2d_hipclaw_options_t *fc2d_hipclaw_get_options(fclaw2d_global_t *glob) { return glob->hipclaw_opt; }

void hipclaw_set_method_parameters(fclaw2d_global_t *glob, const int order[], const int mthlim[],
                                   int mwaves, int use_fwaves)
{
    fc2d_hipclaw_options_t *o = glob->hipclaw_opt;
    int k;
    o->order[0] = order[0]; o->order[1] = order[1];
    for (k = 0; k < mwaves && k < FC2D_HIPCLAW_MWAVES; k++) o->mthlim[k] = mthlim[k];
    o->mwaves = mwaves;
    o->use_fwaves = use_fwaves;
}

/* ------------------------------------------------------------------ buffering */

void hipclaw_store_buffer(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int patchno,
                          int count, int iter, int *slot_buffer)
{
    (void) patchno; (void) count;
    slot_buffer[iter % glob->hipclaw_opt->buffer_len] = patch->slot;
}

double hipclaw_step2_batch(fclaw2d_global_t *glob, const int *slot_buffer, int batch_size,
                           double t, double dt)
{
    double maxcfl = 0.0;
    device_check(hipclaw_step_batch(glob->state, slot_buffer, batch_size, t, dt, &maxcfl),
                 "hipclaw_step_batch");
    return maxcfl;
}

static fclaw2d_patch_t *patch_of_slot(fclaw2d_global_t *glob, int slot, int *blockno, int *patchno)
{
    fclaw2d_domain_t *d = glob->domain;
    int b;
    for (b = d->num_blocks - 1; b >= 0; b--)
        if (slot >= d->blocks[b].patchno_offset) break;
    *blockno = b;
    *patchno = slot - d->blocks[b].patchno_offset;
    return &d->blocks[b].patches[*patchno];
}

/* The single_step_update slot (cudaclaw_update, fc2d_cudaclaw.cpp:272-368):
 * buffer the patch; on a full buffer or the last patch, step the batch on the
 * device and return its CFL number (0 otherwise).  b4step2 is not called, as
 * on the reference's GPU path. */
static double hipclaw_update(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                             int patchno, double t, double dt, void *user)
{
    fclaw2d_single_step_buffer_data_t *bd = (fclaw2d_single_step_buffer_data_t *) user;
    const fc2d_hipclaw_options_t *opt = fc2d_hipclaw_get_options(glob);
    const int len = opt->buffer_len, iter = bd->iter, total = bd->total_count;
    int *buf, n = 0, k;
    double maxcfl = 0.0;
    (void) blockno;
    if (iter == 0) {
        const int size = total < len ? total : len;
        bd->user = malloc(sizeof(int) * (size_t) (size > 0 ? size : 1));
    }
    buf = (int *) bd->user;
    hipclaw_store_buffer(glob, patch, patchno, total, iter, buf);
    if ((iter + 1) % len == 0) n = len;
    else if (iter + 1 == total) n = total % len;
    if (n > 0) {
        maxcfl = hipclaw_step2_batch(glob, buf, n, t, dt);
        if (opt->src_term > 0 && glob->hipclaw_vt->src2) {
            /* source term on the host views of the batch just stepped */
            for (k = 0; k < n; k++) {
                int b, p;
                fclaw2d_patch_t *pt = patch_of_slot(glob, buf[k], &b, &p);
                device_check(hipclaw_get_q(glob->state, glob->qhost + patch_size(glob, glob->clawpatch_opt.meqn) * buf[k], buf[k], 1), "hipclaw_get_q");
                glob->hipclaw_vt->src2(glob, pt, b, p, t, dt);
                device_check(hipclaw_set_q(glob->state, glob->qhost + patch_size(glob, glob->clawpatch_opt.meqn) * buf[k], buf[k], 1), "hipclaw_set_q");
            }
        }
    }
    if (iter == total - 1) {
        free(bd->user);
        bd->user = NULL;
    }
    return maxcfl;
}

/* ------------------------------------------------------------------ core drivers */

double fclaw2d_update_single_step(fclaw2d_global_t *glob, int level, double t, double dt)
{
    fclaw2d_domain_t *d = glob->domain;
    fclaw2d_single_step_buffer_data_t bd;
    double maxcfl = 0.0;
    int b, p, count = 0;
    bd.total_count = 0; bd.iter = 0; bd.user = NULL;
    for (b = 0; b < d->num_blocks; b++)
        for (p = 0; p < d->blocks[b].num_patches; p++)
            if (d->blocks[b].patches[p].level == level) count++;
    bd.total_count = count;
    for (b = 0; b < d->num_blocks; b++)
        for (p = 0; p < d->blocks[b].num_patches; p++) {
            fclaw2d_patch_t *pt = &d->blocks[b].patches[p];
            double cfl;
            if (pt->level != level) continue;
            cfl = glob->patch_vt.single_step_update(glob, pt, b, p, t, dt, &bd);
            bd.iter++;
            glob->count_single_step++;
            maxcfl = fmax(cfl, maxcfl);
        }
    return maxcfl;
}

/* Program numbers: 0 = every level; 1 + (ti - minlevel of the domain) = the
 * levels above a time-interpolated level ti. */
static int program_for(fclaw2d_global_t *glob, int minlevel, int maxlevel, int time_interp,
                       const char *what)
{
    const fclaw2d_domain_t *d = glob->domain;
    if (!time_interp && minlevel <= d->global_minlevel && maxlevel >= d->global_maxlevel) return 0;
    if (time_interp && maxlevel >= d->global_maxlevel && minlevel - 1 >= d->global_minlevel &&
        glob->subcycle)
        return 1 + (minlevel - 1 - d->global_minlevel);
    fprintf(stderr, "%s: level range %d..%d (time_interp %d) is not supported\n", what, minlevel,
            maxlevel, time_interp);
    exit(1);
}

void fclaw2d_ghost_update(fclaw2d_global_t *glob, int minlevel, int maxlevel, double sync_time,
                          int time_interp, int running)
{
    const int k = program_for(glob, minlevel, maxlevel, time_interp, "fclaw2d_ghost_update");
    (void) sync_time; (void) running;
    device_check(hipclaw_ghost_fill_k(glob->state, k), "hipclaw_ghost_fill");
    glob->count_ghost_update++;
}

void fclaw2d_timeinterp(fclaw2d_global_t *glob, int level, double alpha)
{
    fclaw2d_domain_t *d = glob->domain;
    int *slots = (int *) malloc(sizeof(int) * (size_t) (d->local_num_patches + 1));
    int b, p, n = 0;
    for (b = 0; b < d->num_blocks; b++)
        for (p = 0; p < d->blocks[b].num_patches; p++)
            if (d->blocks[b].patches[p].level == level) slots[n++] = d->blocks[b].patches[p].slot;
    device_check(hipclaw_timeinterp(glob->state, slots, n, alpha), "hipclaw_timeinterp");
    free(slots);
}

void fclaw2d_time_sync(fclaw2d_global_t *glob, int minlevel, int maxlevel)
{
    int k;
    if (!glob->time_sync) return;
    k = program_for(glob, minlevel, maxlevel, minlevel > glob->domain->global_minlevel,
                    "fclaw2d_time_sync");
    device_check(hipclaw_time_sync_k(glob->state, k), "hipclaw_time_sync");
}

/* ---- fclaw2d_advance_all_levels with subcycle = T (src/fclaw2d_advance.c:54-246) ---- */
typedef struct {
    int last_step, step_inc, total_steps;
    double dt_step, current_time;
} level_counter_t;

static double advance_level(fclaw2d_global_t *glob, int level, int curr_fine_step, double maxcfl,
                            level_counter_t *tc)
{
    const int minlevel = glob->domain->global_minlevel;
    maxcfl = fmax(maxcfl, fclaw2d_update_single_step(glob, level, tc[level].current_time,
                                                     tc[level].dt_step));
    tc[level].last_step += tc[level].step_inc;
    tc[level].current_time += tc[level].dt_step;
    if (level > minlevel && tc[level - 1].last_step == curr_fine_step) {
        maxcfl = fmax(maxcfl, advance_level(glob, level - 1, tc[level - 1].last_step, maxcfl, tc));
        if (glob->subcycle) {
            /* compute_alpha (:152-170) */
            const double alpha = ((double) (tc[level].last_step % tc[level - 1].step_inc)) /
                                 tc[level - 1].step_inc;
            fclaw2d_timeinterp(glob, level - 1, alpha);
        }
    }
    return maxcfl;
}

static double advance_all_levels_stepwise(fclaw2d_global_t *glob, double t_curr, double dt)
{
    fclaw2d_domain_t *d = glob->domain;
    const int minl = d->global_minlevel, maxl = d->global_maxlevel;
    level_counter_t tc[32];
    double maxcfl = 0.0;
    int level, nf;
    if (maxl >= 32) { fprintf(stderr, "fclaw2d_advance_all_levels: too many levels\n"); exit(1); }
    /* initialize_timestep_counters (:54-110) */
    for (level = minl; level <= maxl; level++) {
        tc[level].last_step = 0;
        tc[level].current_time = t_curr;
        if (glob->subcycle) {
            tc[level].dt_step = dt / (double) (1 << (level - minl));
            tc[level].total_steps = 1 << (level - minl);
            tc[level].step_inc = 1 << (maxl - level);
        } else {
            const int rf = glob->advance_one_step ? 1 : 1 << (maxl - minl);
            tc[level].step_inc = 1;
            tc[level].dt_step = dt / rf;
            tc[level].total_steps = rf;
        }
    }
    for (nf = 0; nf < tc[maxl].total_steps; nf++) {
        int last_step;
        maxcfl = advance_level(glob, maxl, nf, maxcfl, tc);
        last_step = tc[maxl].last_step;
        if (last_step < tc[maxl].total_steps) {
            if (glob->subcycle) {
                int ti = maxl - 1;
                while (last_step % tc[ti].step_inc == 0) ti--;
                fclaw2d_ghost_update(glob, ti + 1, maxl, tc[maxl].current_time, 1, 0);
                if (glob->time_sync) fclaw2d_time_sync(glob, ti + 1, maxl);
            } else {
                fclaw2d_ghost_update(glob, minl, maxl, tc[maxl].current_time, 0, 0);
                if (glob->time_sync) fclaw2d_time_sync(glob, minl, maxl);
            }
        }
    }
    fclaw2d_ghost_update(glob, minl, maxl, tc[maxl].current_time, 0, 0);
    if (glob->time_sync) fclaw2d_time_sync(glob, minl, maxl);
    return maxcfl;
}

/* fclaw2d_advance_all_levels (src/fclaw2d_advance.c:247-357).  Without
 * subcycling every level takes the same steps, so all of them go in one
 * batched launch per step (followed by the device ghost fill and the
 * synchronisation), as one captured sequence. */
double fclaw2d_advance_all_levels(fclaw2d_global_t *glob, double t_curr, double dt)
{
    fclaw2d_domain_t *d = glob->domain;
    double maxcfl = 0.0;
    if (!glob->subcycle && glob->patch_vt.single_step_update == hipclaw_update) {
        const int rf = glob->advance_one_step ? 1 : 1 << (d->global_maxlevel - d->global_minlevel);
        device_check(hipclaw_advance(glob->state, dt / rf, rf, &maxcfl), "hipclaw_advance");
        glob->count_single_step += rf * d->local_num_patches;
        glob->count_ghost_update += rf;
    } else if (glob->subcycle && glob->patch_vt.single_step_update == hipclaw_update) {
        /* the same schedule as advance_all_levels_stepwise, enqueued on the
         * device as one sequence: level steps, time interpolation, partial ghost
         * fills and synchronisations, one Courant number read at the end */
        int b, p;
        device_check(hipclaw_advance_subcycle(glob->state, dt, 1, &maxcfl), "hipclaw_advance_subcycle");
        for (b = 0; b < d->num_blocks; b++)
            for (p = 0; p < d->blocks[b].num_patches; p++)
                glob->count_single_step += 1 << (d->blocks[b].patches[p].level - d->global_minlevel);
        glob->count_ghost_update += 1 << (d->global_maxlevel - d->global_minlevel);
    } else {
        maxcfl = advance_all_levels_stepwise(glob, t_curr, dt);
    }
    glob->curr_time = t_curr + dt;
    glob->curr_dt = dt;
    return maxcfl;
}

/* ------------------------------------------------------------------ solver plugin */

static void hipclaw_qinit_slot(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                               int patchno)
{
    fc2d_hipclaw_qinit(glob, patch, blockno, patchno);
}

static void hipclaw_setaux_slot(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                                int patchno)
{
    fc2d_hipclaw_setaux(glob, patch, blockno, patchno);
}

void fc2d_hipclaw_solver_initialize(fclaw2d_global_t *glob)
{
    fc2d_hipclaw_vtable_t *vt = (fc2d_hipclaw_vtable_t *) calloc(1, sizeof(*vt));
    if (glob->hipclaw_vt) {
        fprintf(stderr, "fc2d_hipclaw_solver_initialize: called twice\n");
        exit(1);
    }
    vt->riemann_solver = HIPCLAW_RP_SHALLOW;
    vt->is_set = 1;
    glob->hipclaw_vt = vt;
    glob->patch_vt.initialize = hipclaw_qinit_slot;
    glob->patch_vt.setup = hipclaw_setaux_slot;
    glob->patch_vt.physical_bc = fc2d_hipclaw_bc2;
    glob->patch_vt.single_step_update = hipclaw_update;
    glob->patch_vt.is_set = 1;
}

fc2d_hipclaw_vtable_t *fc2d_hipclaw_vt(fclaw2d_global_t *glob)
{
    if (!glob->hipclaw_vt || !glob->hipclaw_vt->is_set) {
        fprintf(stderr, "fc2d_hipclaw_vt: solver not initialized\n");
        exit(1);
    }
    return glob->hipclaw_vt;
}

void fc2d_hipclaw_initialize_GPUs(fclaw2d_global_t *glob)
{
    char name[256];
    (void) glob;
    device_check(hipclaw_initialize_device(0, name, sizeof(name)), "hipclaw_initialize_device");
}

/* Ghost-fill and time-sync programs of the current domain.  The op buffer
 * is sized for the usual forest (at most 16 ghost and 8 sync ops per leaf),
 * so each program is built once; a larger one is built again into a buffer
 * of the size the first call reported. */
static void install_programs(fclaw2d_global_t *glob)
{
    const fc2d_hipclaw_options_t *opt = fc2d_hipclaw_get_options(glob);
    fclaw2d_domain_t *d = glob->domain;
    const fclaw2d_clawpatch_options_t *o = &glob->clawpatch_opt;
    int *ops, phase[7], nops, cap = 16 * fclaw2d_hip_forest_num_patches(d->forest) + 16;
    ops = (int *) malloc(sizeof(int) * 8 * (size_t) cap);
    if (!ops) { fprintf(stderr, "install_programs: out of memory\n"); exit(1); }
#define GROW_OPS(n) do {                                                              \
        if ((n) > cap) {                                                              \
            free(ops);                                                                \
            cap = (n);                                                                \
            ops = (int *) malloc(sizeof(int) * 8 * (size_t) cap);                     \
            if (!ops) { fprintf(stderr, "install_programs: out of memory\n"); exit(1); } \
        }                                                                             \
    } while (0)
    nops = fclaw2d_hip_forest_ghost_program(d->forest, o->mx, o->my, o->mbc, opt->mthbc, ops, cap, phase);
    if (nops > cap) {
        GROW_OPS(nops);
        fclaw2d_hip_forest_ghost_program(d->forest, o->mx, o->my, o->mbc, opt->mthbc, ops, cap, phase);
    }
    device_check(hipclaw_set_ghost_program(glob->state, ops, nops, phase), "hipclaw_set_ghost_program");
    if (glob->subcycle) {
        /* one program per time-interpolated level ti: levels ti+1..maxlevel */
        int ti;
        if (d->global_maxlevel - d->global_minlevel >= HIPCLAW_MAX_PROGRAMS) {
            fprintf(stderr, "fc2d_hipclaw_allocate_buffers: too many levels\n");
            exit(1);
        }
        for (ti = d->global_minlevel; ti < d->global_maxlevel; ti++) {
            nops = fclaw2d_hip_forest_ghost_program_range(d->forest, o->mx, o->my, o->mbc, opt->mthbc,
                                                          ti + 1, d->global_maxlevel, 1, ops, cap, phase);
            if (nops > cap) {
                GROW_OPS(nops);
                fclaw2d_hip_forest_ghost_program_range(d->forest, o->mx, o->my, o->mbc, opt->mthbc, ti + 1,
                                                       d->global_maxlevel, 1, ops, cap, phase);
            }
            device_check(hipclaw_set_ghost_program_k(glob->state, 1 + ti - d->global_minlevel, ops,
                                                     nops, phase), "hipclaw_set_ghost_program");
        }
    }
    if (glob->subcycle) {
        /* the level of every slot, for the device-resident subcycled schedule */
        int b, p, *level = (int *) malloc(sizeof(int) * (size_t) (d->local_num_patches + 1));
        for (b = 0; b < d->num_blocks; b++)
            for (p = 0; p < d->blocks[b].num_patches; p++)
                level[d->blocks[b].patches[p].slot] = d->blocks[b].patches[p].level;
        device_check(hipclaw_set_levels(glob->state, level, d->local_num_patches), "hipclaw_set_levels");
        free(level);
    }
    if (glob->time_sync) {
        /* program 0: every level; with subcycling also one per level range */
        int minl;
        int rcap = 4 * fclaw2d_hip_forest_num_patches(d->forest) + 4;   /* one per face */
        int *res = (int *) malloc(sizeof(int) * 2 * (size_t) rcap);
        if (!res) { fprintf(stderr, "install_programs: out of memory\n"); exit(1); }
        for (minl = d->global_minlevel; minl <= (glob->subcycle ? d->global_maxlevel : d->global_minlevel); minl++) {
            int nres = 0;
            nops = fclaw2d_hip_forest_sync_program_range(d->forest, o->mx, o->my, minl, ops, cap, res, rcap, &nres);
            if (nops < 0) {
                fprintf(stderr, "fc2d_hipclaw_allocate_buffers: time sync needs mx = my\n");
                exit(1);
            }
            if (nops > cap || nres > rcap) {
                GROW_OPS(nops);
                if (nres > rcap) {
                    free(res);
                    rcap = nres;
                    res = (int *) malloc(sizeof(int) * 2 * (size_t) rcap);
                    if (!res) { fprintf(stderr, "install_programs: out of memory\n"); exit(1); }
                }
                fclaw2d_hip_forest_sync_program_range(d->forest, o->mx, o->my, minl, ops, cap, res, rcap, &nres);
            }
            /* without subcycling every register is rewritten by every update */
            device_check(hipclaw_set_sync_program_k(glob->state, minl == d->global_minlevel ? 0 : minl - d->global_minlevel,
                                                    ops, nops, res, glob->subcycle ? nres : 0),
                         "hipclaw_set_sync_program");
        }
        free(res);
    }
#undef GROW_OPS
    free(ops);
}

/* Host views of q and aux for every patch (zeroed). */
static void host_views(fclaw2d_global_t *glob)
{
    fclaw2d_domain_t *d = glob->domain;
    const fclaw2d_clawpatch_options_t *o = &glob->clawpatch_opt;
    free(glob->qhost);
    free(glob->auxhost);
    glob->auxhost = NULL;
    glob->qhost = (double *) calloc(patch_size(glob, o->meqn) * d->local_num_patches, sizeof(double));
    if (o->maux > 0)
        glob->auxhost = (double *) calloc(patch_size(glob, o->maux) * d->local_num_patches, sizeof(double));
}

/* Mesh widths of every patch, by slot. */
static void patch_widths(fclaw2d_global_t *glob, double *dx, double *dy)
{
    fclaw2d_domain_t *d = glob->domain;
    int b, p;
    for (b = 0; b < d->num_blocks; b++)
        for (p = 0; p < d->blocks[b].num_patches; p++) {
            fclaw2d_patch_t *pt = &d->blocks[b].patches[p];
            int mx, my, mbc;
            double xl, yl;
            fclaw2d_clawpatch_grid_data(glob, pt, &mx, &my, &mbc, &xl, &yl, &dx[pt->slot], &dy[pt->slot]);
        }
}

void fc2d_hipclaw_allocate_buffers(fclaw2d_global_t *glob)
{
    const fc2d_hipclaw_options_t *opt = fc2d_hipclaw_get_options(glob);
    const fc2d_hipclaw_vtable_t *vt = fc2d_hipclaw_vt(glob);
    fclaw2d_domain_t *d = glob->domain;
    const fclaw2d_clawpatch_options_t *o = &glob->clawpatch_opt;
    hipclaw_config_t cfg;
    double *dx, *dy;
    int k;
    if (!opt || !d) {
        fprintf(stderr, "fc2d_hipclaw_allocate_buffers: options or domain missing\n");
        exit(1);
    }
    if (fc2d_hipclaw_check(opt)) exit(1);
    fc2d_hipclaw_deallocate_buffers(glob);
    memset(&cfg, 0, sizeof(cfg));
    cfg.solver = vt->riemann_solver;
    if (vt->rp_solver)   /* the application's own solver, as cuda_rpn2 / cuda_rpt2 */
        device_check(hipclaw_register_solver(vt->rp_solver, &cfg.solver), "hipclaw_register_solver");
    cfg.precision = opt->precision;
    cfg.mx = o->mx; cfg.my = o->my; cfg.mbc = o->mbc;
    cfg.meqn = o->meqn; cfg.maux = o->maux; cfg.mwaves = opt->mwaves;
    cfg.order[0] = opt->order[0]; cfg.order[1] = opt->order[1];
    for (k = 0; k < FC2D_HIPCLAW_MWAVES; k++) cfg.mthlim[k] = k < opt->mwaves ? opt->mthlim[k] : 0;
    for (k = 0; k < 4; k++) cfg.mthbc[k] = opt->mthbc[k];
    cfg.grav = opt->grav; cfg.gamma = opt->gamma; cfg.tile = opt->tile;
    for (k = 0; k < 8; k++) cfg.rp_params[k] = opt->rp_params[k];
    cfg.time_sync = glob->time_sync;
    cfg.subcycle = glob->subcycle;
    device_check(hipclaw_create(&cfg, d->local_num_patches, &glob->state), "hipclaw_create");
    dx = (double *) malloc(sizeof(double) * d->local_num_patches);
    dy = (double *) malloc(sizeof(double) * d->local_num_patches);
    patch_widths(glob, dx, dy);
    device_check(hipclaw_set_geometry(glob->state, d->local_num_patches, dx, dy), "hipclaw_set_geometry");
    free(dx); free(dy);
    install_programs(glob);
    host_views(glob);
}

void fc2d_hipclaw_deallocate_buffers(fclaw2d_global_t *glob)
{
    if (glob->state) hipclaw_destroy(glob->state);
    glob->state = NULL;
    free(glob->qhost); glob->qhost = NULL;
    free(glob->auxhost); glob->auxhost = NULL;
}

void fc2d_hipclaw_upload(fclaw2d_global_t *glob)
{
    const int n = glob->domain->local_num_patches;
    device_check(hipclaw_set_q(glob->state, glob->qhost, 0, n), "hipclaw_set_q");
    if (glob->auxhost) device_check(hipclaw_set_aux(glob->state, glob->auxhost, 0, n), "hipclaw_set_aux");
}

void fc2d_hipclaw_download(fclaw2d_global_t *glob)
{
    device_check(hipclaw_get_q(glob->state, glob->qhost, 0, glob->domain->local_num_patches),
                 "hipclaw_get_q");
}

void fc2d_hipclaw_setprob(fclaw2d_global_t *glob)
{
    fc2d_hipclaw_vtable_t *vt = fc2d_hipclaw_vt(glob);
    if (vt->fort_setprob) vt->fort_setprob();
}

void fc2d_hipclaw_qinit(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno, int patchno)
{
    fc2d_hipclaw_vtable_t *vt = fc2d_hipclaw_vt(glob);
    int mx, my, mbc, meqn, maux;
    double xl, yl, dx, dy, *q, *aux;
    (void) blockno; (void) patchno;
    if (!vt->fort_qinit) return;
    fclaw2d_clawpatch_grid_data(glob, patch, &mx, &my, &mbc, &xl, &yl, &dx, &dy);
    fclaw2d_clawpatch_soln_data(glob, patch, &q, &meqn);
    fclaw2d_clawpatch_aux_data(glob, patch, &aux, &maux);
    vt->fort_qinit(&mx, &my, &meqn, &mbc, &mx, &my, &xl, &yl, &dx, &dy, q, &maux, aux);
}

void fc2d_hipclaw_setaux(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno, int patchno)
{
    fc2d_hipclaw_vtable_t *vt = fc2d_hipclaw_vt(glob);
    int mx, my, mbc, maux;
    double xl, yl, dx, dy, *aux;
    (void) blockno; (void) patchno;
    if (!vt->fort_setaux) return;
    fclaw2d_clawpatch_grid_data(glob, patch, &mx, &my, &mbc, &xl, &yl, &dx, &dy);
    fclaw2d_clawpatch_aux_data(glob, patch, &aux, &maux);
    vt->fort_setaux(&mx, &my, &mbc, &mx, &my, &xl, &yl, &dx, &dy, &maux, aux);
}

void fc2d_hipclaw_b4step2(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                          int patchno, double t, double dt)
{
    fc2d_hipclaw_vtable_t *vt = fc2d_hipclaw_vt(glob);
    if (vt->b4step2) vt->b4step2(glob, patch, blockno, patchno, t, dt);
}

/* Physical BC on a patch's host view (clawpack46_bc2_default semantics, or the
 * user's fort_bc2); the device path applies the same BC inside the ghost fill. */
void fc2d_hipclaw_bc2(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno, int patchno,
                      double t, double dt, int intersects_bc[], int time_interp)
{
    fc2d_hipclaw_vtable_t *vt = fc2d_hipclaw_vt(glob);
    const fc2d_hipclaw_options_t *opt = fc2d_hipclaw_get_options(glob);
    int mx, my, mbc, meqn, maux, mthbc[4], k, m, i, j, b;
    double xl, yl, dx, dy, *q, *aux;
    (void) blockno; (void) patchno; (void) time_interp;
    fclaw2d_clawpatch_grid_data(glob, patch, &mx, &my, &mbc, &xl, &yl, &dx, &dy);
    fclaw2d_clawpatch_soln_data(glob, patch, &q, &meqn);
    fclaw2d_clawpatch_aux_data(glob, patch, &aux, &maux);
    for (k = 0; k < 4; k++) mthbc[k] = intersects_bc[k] ? opt->mthbc[k] : -1;
    if (vt->fort_bc2) {
        vt->fort_bc2(&mx, &my, &meqn, &mbc, &mx, &my, &xl, &yl, &dx, &dy, q, &maux, aux, &t, &dt, mthbc);
        return;
    }
#define QV(i, j, m) q[((size_t) (m) * (my + 2 * mbc) + (j) + mbc - 1) * (mx + 2 * mbc) + (i) + mbc - 1]
    for (k = 0; k < 4; k++) {
        if (mthbc[k] != 1 && mthbc[k] != 3) continue;
        for (m = 0; m < meqn; m++)
            for (b = 1; b <= mbc; b++) {
                if (k < 2)
                    for (j = 1 - mbc; j <= my + mbc; j++) {
                        const int id = k == 0 ? 1 - b : mx + b;
                        const int is = mthbc[k] == 1 ? (k == 0 ? 1 : mx) : (k == 0 ? b : mx + 1 - b);
                        QV(id, j, m) = (mthbc[k] == 3 && m == 1) ? -QV(is, j, m) : QV(is, j, m);
                    }
                else
                    for (i = 1 - mbc; i <= mx + mbc; i++) {
                        const int jd = k == 2 ? 1 - b : my + b;
                        const int js = mthbc[k] == 1 ? (k == 2 ? 1 : my) : (k == 2 ? b : my + 1 - b);
                        QV(i, jd, m) = (mthbc[k] == 3 && m == 2) ? -QV(i, js, m) : QV(i, js, m);
                    }
            }
    }
#undef QV
}

void fc2d_hipclaw_src2(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno, int patchno,
                       double t, double dt)
{
    fc2d_hipclaw_vtable_t *vt = fc2d_hipclaw_vt(glob);
    int mx, my, mbc, meqn, maux;
    double xl, yl, dx, dy, *q, *aux;
    (void) blockno; (void) patchno;
    if (!vt->fort_src2) return;
    fclaw2d_clawpatch_grid_data(glob, patch, &mx, &my, &mbc, &xl, &yl, &dx, &dy);
    fclaw2d_clawpatch_soln_data(glob, patch, &q, &meqn);
    fclaw2d_clawpatch_aux_data(glob, patch, &aux, &maux);
    vt->fort_src2(&mx, &my, &meqn, &mbc, &mx, &my, &xl, &yl, &dx, &dy, q, &maux, aux, &t, &dt);
}

/* ------------------------------------------------------------------ ASCII output */

/* A value in Fortran Ew.d edit form (0.d1..dd E+xx, or +xxx for three-digit
 * exponents), right-justified in w characters. */
static void fortran_e(char *out, double x, int w, int d)
{
    char digits[64], body[96];
    int e10 = 0, neg = x < 0.0, len;
    if (x == 0.0) {
        memset(digits, '0', (size_t) d);
        digits[d] = 0;
        e10 = 0;
    } else {
        char tmp[96];
        char *p, *q;
        snprintf(tmp, sizeof tmp, "%.*e", d - 1, neg ? -x : x);  /* d.ddd..e+XX */
        q = digits;
        for (p = tmp; *p && *p != 'e'; p++)
            if (*p >= '0' && *p <= '9') *q++ = *p;
        *q = 0;
        e10 = atoi(p + 1) + 1;
    }
    if (e10 >= -99 && e10 <= 99)
        len = snprintf(body, sizeof body, "%s0.%sE%c%02d", neg ? "-" : "", digits,
                       e10 < 0 ? '-' : '+', e10 < 0 ? -e10 : e10);
    else
        len = snprintf(body, sizeof body, "%s0.%s%c%03d", neg ? "-" : "", digits,
                       e10 < 0 ? '-' : '+', e10 < 0 ? -e10 : e10);
    if (len >= w) {
        memcpy(out, body, (size_t) len + 1);
    } else {
        memset(out, ' ', (size_t) (w - len));
        memcpy(out + (w - len), body, (size_t) len + 1);
    }
}

int fclaw2d_clawpatch_time_header_ascii_files(const char *fname_q, const char *fname_t,
                                              double time, int meqn, int maux, int ngrids)
{
    char e[64];
    FILE *f = fopen(fname_t, "w");
    if (!f) return -1;
    fortran_e(e, time, 30, 20);
    fprintf(f, "%s    time\n%5d                 meqn\n%5d                 ngrids\n"
               "%5d                 num_aux\n%5d                 num_dim\n",
            e, meqn, ngrids, maux, 2);
    fclose(f);
    f = fopen(fname_q, "w");
    if (!f) return -1;
    fclose(f);
    return 0;
}

int fclaw2d_clawpatch_output_ascii_patch(const char *fname_q, int mx, int my, int meqn, int mbc,
                                         double xlower, double ylower, double dx, double dy,
                                         const double *q, int patch_num, int level, int blockno,
                                         int mpirank)
{
    const int nx = mx + 2 * mbc, ny = my + 2 * mbc;
    char e[64];
    int i, j, m;
    FILE *f;
    if (meqn > 20) return -1;
    f = fopen(fname_q, "a");
    if (!f) return -1;
    fprintf(f, "%5d                 grid_number\n%5d                 AMR_level\n"
               "%5d                 block_number\n%5d                 mpi_rank\n"
               "%5d                 mx\n%5d                 my\n",
            patch_num, level, blockno, mpirank, mx, my);
    fortran_e(e, xlower, 24, 16); fprintf(f, "%s    xlow\n", e);
    fortran_e(e, ylower, 24, 16); fprintf(f, "%s    ylow\n", e);
    fortran_e(e, dx, 24, 16); fprintf(f, "%s    dx\n", e);
    fortran_e(e, dy, 24, 16); fprintf(f, "%s    dy\n\n", e);
    for (j = 1; j <= my; j++) {
        for (i = 1; i <= mx; i++) {
            for (m = 0; m < meqn; m++) {
                double v = q[((size_t) m * ny + (size_t) (j + mbc - 1)) * nx + (size_t) (i + mbc - 1)];
                if (fabs(v) < 1e-99) v = 0.0;  /* as the reference writer does */
                fortran_e(e, v, 26, 16);
                fputs(e, f);
            }
            fputc('\n', f);
        }
        fputs("  \n", f);
    }
    fclose(f);
    return 0;
}

/* fclaw2d_clawpatch_output_ascii (src/patches/clawpatch/fclaw2d_clawpatch_output_ascii.c:139-160)
 * with the writers of fclaw2d_clawpatch46_output_ascii.f: fort.tXXXX and
 * fort.qXXXX in the current directory, from the device-resident solution. */
void fclaw2d_clawpatch_output_ascii(fclaw2d_global_t *glob, int iframe)
{
    fclaw2d_domain_t *d = glob->domain;
    const fclaw2d_clawpatch_options_t *o = &glob->clawpatch_opt;
    char fq[32], ft[32];
    int b, p;
    fc2d_hipclaw_download(glob);
    snprintf(fq, sizeof fq, "fort.q%04d", iframe);
    snprintf(ft, sizeof ft, "fort.t%04d", iframe);
    if (fclaw2d_clawpatch_time_header_ascii_files(fq, ft, glob->curr_time, o->meqn, o->maux,
                                                  d->local_num_patches)) {
        fprintf(stderr, "fclaw2d_clawpatch_output_ascii: cannot write %s\n", ft);
        exit(1);
    }
    for (b = 0; b < d->num_blocks; b++)
        for (p = 0; p < d->blocks[b].num_patches; p++) {
            fclaw2d_patch_t *pt = &d->blocks[b].patches[p];
            int mx, my, mbc;
            double xl, yl, dx, dy;
            fclaw2d_clawpatch_grid_data(glob, pt, &mx, &my, &mbc, &xl, &yl, &dx, &dy);
            if (fclaw2d_clawpatch_output_ascii_patch(fq, mx, my, o->meqn, mbc, xl, yl, dx, dy,
                                                     glob->qhost + patch_size(glob, o->meqn) * (size_t) pt->slot,
                                                     d->blocks[b].patchno_offset + p, pt->level, b, 0)) {
                fprintf(stderr, "fclaw2d_clawpatch_output_ascii: cannot write %s\n", fq);
                exit(1);
            }
        }
}

/* ------------------------------------------------------------------ regrid */

/* Replace the domain by one over the (adapted) leaves of its forest. */
static void rebuild_domain(fclaw2d_global_t *glob)
{
    fclaw2d_domain_t *d = glob->domain;
    fclaw2d_domain_t *nd = fclaw2d_domain_new_from_forest(d->forest, d->nbx, d->nby, d->periodic_x,
                                                          d->periodic_y, d->ax, d->bx, d->ay, d->by);
    if (!nd) {
        fprintf(stderr, "fclaw2d_regrid: out of memory\n");
        exit(1);
    }
    d->forest = NULL;   /* now owned by the new domain */
    fclaw2d_domain_destroy(d);
    glob->domain = nd;
}

/* Tag every patch on the device; refine[]/coarsen[] by leaf (= slot).
 * Coarsening is never asked for while the mesh is being initialized
 * (fclaw2d_clawpatch46_tag4coarsening.f: initflag). */
static void tag_leaves(fclaw2d_global_t *glob, int domain_init, int *refine, int *coarsen)
{
    const int n = glob->domain->local_num_patches;
    int p, *ec = (int *) malloc(sizeof(int) * (size_t) (n + 1));
    if (glob->refinement_criteria != FCLAW_REFINE_CRITERIA_GRADIENT) {
        fprintf(stderr, "fclaw2d_regrid: refinement-criteria %d is not supported; the device tags "
                        "with the gradient criterion only (refinement_criteria = %d)\n",
                glob->refinement_criteria, FCLAW_REFINE_CRITERIA_GRADIENT);
        exit(1);
    }
    if (glob->threshold_variable < 1 || glob->threshold_variable > glob->clawpatch_opt.meqn) {
        fprintf(stderr, "fclaw2d_regrid: threshold_variable %d is not in 1..meqn\n",
                glob->threshold_variable);
        exit(1);
    }
    device_check(hipclaw_tag_patches(glob->state, glob->threshold_variable - 1,
                                     glob->refine_threshold, glob->coarsen_threshold, refine, ec),
                 "hipclaw_tag_patches");
    for (p = 0; p < n; p++) coarsen[p] = !domain_init && !ec[p];
    free(ec);
}

/* Adapt the forest; returns 1 if the leaves changed.  kind/src receive the
 * provenance (caller frees). */
static int adapt_forest(fclaw2d_global_t *glob, const int *refine, const int *coarsen, int **kind,
                        int **src)
{
    fclaw2d_domain_t *d = glob->domain;
    const int n0 = d->local_num_patches;
    int p, changed, n = fclaw2d_hip_forest_adapt(d->forest, refine, coarsen, glob->minlevel,
                                                 glob->maxlevel);
    if (n <= 0) {
        fprintf(stderr, "fclaw2d_regrid: forest adapt failed\n");
        exit(1);
    }
    *kind = (int *) malloc(sizeof(int) * (size_t) n);
    *src = (int *) malloc(sizeof(int) * 4 * (size_t) n);
    if (fclaw2d_hip_forest_provenance(d->forest, *kind, *src) != n) {
        fprintf(stderr, "fclaw2d_regrid: forest provenance failed\n");
        exit(1);
    }
    changed = n != n0;
    for (p = 0; p < n && !changed; p++)
        changed = (*kind)[p] != FCLAW2D_HIP_REGRID_SAME || (*src)[4 * p] != p;
    return changed;
}

/* setaux (patch setup) then qinit on every patch's host view, uploaded. */
static void initialize_patches(fclaw2d_global_t *glob, int with_q)
{
    fclaw2d_domain_t *d = glob->domain;
    int b, p;
    for (b = 0; b < d->num_blocks; b++)
        for (p = 0; p < d->blocks[b].num_patches; p++) {
            fclaw2d_patch_t *pt = &d->blocks[b].patches[p];
            if (glob->patch_vt.setup) glob->patch_vt.setup(glob, pt, b, p);
            if (with_q && glob->patch_vt.initialize) glob->patch_vt.initialize(glob, pt, b, p);
        }
    if (with_q)
        device_check(hipclaw_set_q(glob->state, glob->qhost, 0, d->local_num_patches), "hipclaw_set_q");
    if (glob->auxhost)
        device_check(hipclaw_set_aux(glob->state, glob->auxhost, 0, d->local_num_patches),
                     "hipclaw_set_aux");
}

void fclaw2d_initialize(fclaw2d_global_t *glob)
{
    int level;
    fc2d_hipclaw_allocate_buffers(glob);
    initialize_patches(glob, 1);
    fclaw2d_ghost_update(glob, glob->domain->global_minlevel, glob->domain->global_maxlevel, 0.0, 0, 0);
    for (level = glob->minlevel; level < glob->maxlevel; level++) {
        const int n = glob->domain->local_num_patches;
        int *refine = (int *) malloc(sizeof(int) * (size_t) (n + 1));
        int *coarsen = (int *) malloc(sizeof(int) * (size_t) (n + 1));
        int *kind, *src, changed;
        tag_leaves(glob, 1, refine, coarsen);
        changed = adapt_forest(glob, refine, coarsen, &kind, &src);
        free(refine); free(coarsen); free(kind); free(src);
        if (!changed) break;   /* no new initial refinement */
        rebuild_domain(glob);
        /* new patches are initialized, not interpolated (domain_init) */
        fc2d_hipclaw_allocate_buffers(glob);
        initialize_patches(glob, 1);
        fclaw2d_ghost_update(glob, glob->domain->global_minlevel, glob->domain->global_maxlevel,
                             0.0, 0, 0);
    }
    glob->curr_time = 0.0;
}

void fclaw2d_regrid(fclaw2d_global_t *glob)
{
    const int n = glob->domain->local_num_patches;
    int *refine = (int *) malloc(sizeof(int) * (size_t) (n + 1));
    int *coarsen = (int *) malloc(sizeof(int) * (size_t) (n + 1));
    int *kind, *src, changed;
    tag_leaves(glob, 0, refine, coarsen);
    changed = adapt_forest(glob, refine, coarsen, &kind, &src);
    free(refine); free(coarsen);
    if (changed) {
        fclaw2d_domain_t *d;
        double *dx, *dy;
        rebuild_domain(glob);
        d = glob->domain;
        dx = (double *) malloc(sizeof(double) * (size_t) d->local_num_patches);
        dy = (double *) malloc(sizeof(double) * (size_t) d->local_num_patches);
        patch_widths(glob, dx, dy);
        device_check(hipclaw_regrid(glob->state, d->local_num_patches, kind, src, dx, dy),
                     "hipclaw_regrid");
        free(dx); free(dy);
        install_programs(glob);
        host_views(glob);
        initialize_patches(glob, 0);
        /* every level is at the same time when the mesh changes */
        fclaw2d_ghost_update(glob, d->global_minlevel, d->global_maxlevel, glob->curr_time, 0, 0);
        glob->count_amr_new_domain++;
    }
    free(kind); free(src);
    glob->count_amr_regrid++;
}

/* ------------------------------------------------------------------ run loop */

void fclaw2d_save_time_step(fclaw2d_global_t *glob)
{
    device_check(hipclaw_save_step(glob->state), "hipclaw_save_step");
}

void fclaw2d_restore_time_step(fclaw2d_global_t *glob)
{
    device_check(hipclaw_restore_step(glob->state), "hipclaw_restore_step");
}

void fclaw2d_diagnostics_gather(fclaw2d_global_t *glob, int init_flag)
{
    double hi[4], lo[4];
    int m;
    const int meqn = glob->clawpatch_opt.meqn;
    if (!glob->conservation_check) return;
    if (meqn > 4) { fprintf(stderr, "fclaw2d_diagnostics_gather: meqn > 4\n"); exit(1); }
    device_check(hipclaw_mass(glob->state, hi, lo), "hipclaw_mass");
    for (m = 0; m < meqn; m++) {
        glob->cons_mass[m] = hi[m] + lo[m];
        if (init_flag) glob->cons_mass0[m] = glob->cons_mass[m];
        printf("sum[%d] =  %24.16e  %24.16e\n", m, glob->cons_mass[m],
               fabs(glob->cons_mass[m] - glob->cons_mass0[m]));
    }
    fflush(stdout);
}

/* fclaw2d_output_frame: the frame files the solver's options ask for */
static void output_frame(fclaw2d_global_t *glob, int iframe)
{
    const fc2d_hipclaw_options_t *opt = glob->hipclaw_opt;
    if (!glob->output || !opt) return;
    if (opt->ascii_out) fclaw2d_clawpatch_output_ascii(glob, iframe);
    if (opt->vtk_out) fclaw2d_clawpatch_output_vtk(glob, iframe);
}

static int level_factor(const fclaw2d_global_t *glob)
{
    return 1 << (glob->maxlevel - glob->minlevel);
}

/* A step whose Courant number is above max_cfl is rejected: the saved state
 * comes back and dt shrinks to give desired_cfl (src/fclaw2d_run.c:220-232,
 * 379-395).  Returns 1 if the step stands. */
static int accept_step(fclaw2d_global_t *glob, double maxcfl, double *dt_minlevel)
{
    if (glob->reduce_cfl && maxcfl > glob->max_cfl && !glob->use_fixed_dt) {
        fclaw2d_restore_time_step(glob);
        *dt_minlevel = *dt_minlevel * glob->desired_cfl / maxcfl;
        glob->count_retake++;
        return 0;
    }
    return 1;
}

/* Output times 0, T/nout, 2T/nout, ..., T (src/fclaw2d_run.c:120-300) */
static void run_outstyle_1(fclaw2d_global_t *glob)
{
    const double dt_outer = glob->tfinal / (double) glob->nout;
    double dt_minlevel = glob->initial_dt, t_curr = 0.0;
    int n, n_inner = 0, iframe = 0;
    fclaw2d_diagnostics_gather(glob, 1);
    output_frame(glob, iframe);
    for (n = 0; n < glob->nout; n++) {
        const double tend = t_curr + dt_outer;
        glob->curr_time = t_curr;
        while (t_curr < tend) {
            double dt_step = dt_minlevel, maxcfl;
            int took_small_step = 0;
            if (!glob->use_fixed_dt) fclaw2d_save_time_step(glob);
            if (glob->advance_one_step) dt_step /= level_factor(glob);
            if (!glob->use_fixed_dt) {
                /* a step at most 1 % longer to land on tend, rather than a tiny one after it */
                const double small_step = tend - (t_curr + dt_step);
                if (small_step < 1e-2 * dt_step) {
                    dt_step = tend - t_curr;
                    took_small_step = small_step < 0;
                }
            }
            glob->curr_dt = dt_step;
            maxcfl = fclaw2d_advance_all_levels(glob, t_curr, dt_step);
            if (!accept_step(glob, maxcfl, &dt_minlevel)) continue;
            n_inner++;
            t_curr += dt_step;
            /* a step shortened to land on tend keeps the dt it would have taken */
            if (!glob->use_fixed_dt && !took_small_step)
                dt_minlevel = dt_minlevel * glob->desired_cfl / maxcfl;
            glob->curr_time = t_curr;
            if (glob->advance_one_step) fclaw2d_diagnostics_gather(glob, 0);
            if (glob->regrid_interval > 0 && n_inner % glob->regrid_interval == 0) fclaw2d_regrid(glob);
        }
        fclaw2d_diagnostics_gather(glob, 0);
        glob->curr_time = t_curr;
        output_frame(glob, ++iframe);
    }
}

/* nout steps, a frame every nstep steps (src/fclaw2d_run.c:306-427) */
static void run_outstyle_3(fclaw2d_global_t *glob)
{
    const int lf = level_factor(glob);
    const int per_level = !glob->subcycle && glob->advance_one_step;
    double dt_minlevel = glob->initial_dt, t_curr = 0.0;
    int nstep_outer = glob->nout, nstep_inner = glob->nstep, n = 0, iframe = 0;
    fclaw2d_diagnostics_gather(glob, 1);
    output_frame(glob, iframe);
    glob->curr_time = t_curr;
    if (per_level && !glob->outstyle_uses_maxlevel) {
        nstep_outer *= lf;
        nstep_inner *= lf;
    }
    while (n < nstep_outer) {
        double dt_step = per_level ? dt_minlevel / lf : dt_minlevel, maxcfl;
        if (!glob->use_fixed_dt) fclaw2d_save_time_step(glob);
        glob->curr_dt = dt_step;
        maxcfl = fclaw2d_advance_all_levels(glob, t_curr, dt_step);
        if (!accept_step(glob, maxcfl, &dt_minlevel)) continue;
        t_curr += dt_step;
        glob->curr_time = t_curr;
        if (!glob->use_fixed_dt) dt_minlevel = dt_minlevel * glob->desired_cfl / maxcfl;
        n++;
        if (glob->regrid_interval > 0 && n % glob->regrid_interval == 0) fclaw2d_regrid(glob);
        if (n % nstep_inner == 0) {
            fclaw2d_diagnostics_gather(glob, 0);
            output_frame(glob, ++iframe);
        }
    }
}

/* nout steps of the fixed dt initial_dt, no CFL check (src/fclaw2d_run.c:429-514) */
static void run_outstyle_4(fclaw2d_global_t *glob)
{
    double t_curr = 0.0;
    int n = 0, iframe = 0;
    output_frame(glob, iframe);
    fclaw2d_diagnostics_gather(glob, 1);
    glob->curr_time = t_curr;
    while (n < glob->nout) {
        fclaw2d_advance_all_levels(glob, t_curr, glob->initial_dt);
        t_curr += glob->initial_dt;
        n++;
        glob->curr_time = t_curr;
        if (glob->regrid_interval > 0 && n % glob->regrid_interval == 0) fclaw2d_regrid(glob);
        if (n % glob->nstep == 0) {
            fclaw2d_diagnostics_gather(glob, 0);
            output_frame(glob, ++iframe);
        }
    }
}

void fclaw2d_run(fclaw2d_global_t *glob)
{
    switch (glob->outstyle) {
    case 0:
        output_frame(glob, 0);
        fclaw2d_diagnostics_gather(glob, 1);
        break;
    case 1: run_outstyle_1(glob); break;
    case 3: run_outstyle_3(glob); break;
    case 4: run_outstyle_4(glob); break;
    default:
        fprintf(stderr, "Outstyle %d not implemented yet\n", glob->outstyle);
        exit(0);
    }
}

/* ------------------------------------------------------------------ VTK output */

/* The byte layout of a frame (UnstructuredGrid, raw appended data with UInt64
 * byte counts, quads of type 9) is the reference's, so that ParaView/VisIt
 * tooling written for ForestClaw reads these files unchanged. */

void fclaw2d_output_vtk_coordinate_cb(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                                      int patchno, char *a)
{
    int mx, my, mbc, i, j;
    double xl, yl, dx, dy, *d = (double *) a;
    (void) blockno; (void) patchno;
    fclaw2d_clawpatch_grid_data(glob, patch, &mx, &my, &mbc, &xl, &yl, &dx, &dy);
    for (j = 0; j <= my; j++)
        for (i = 0; i <= mx; i++) {
            *d++ = xl + i * dx;
            *d++ = yl + j * dy;
            *d++ = 0;
        }
}

void fclaw2d_output_vtk_value_cb(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                                 int patchno, char *a)
{
    int mx, my, mbc, meqn, i, j, k;
    double xl, yl, dx, dy, *q;
    float *f = (float *) a;
    (void) blockno; (void) patchno;
    fclaw2d_clawpatch_grid_data(glob, patch, &mx, &my, &mbc, &xl, &yl, &dx, &dy);
    fclaw2d_clawpatch_soln_data(glob, patch, &q, &meqn);
    for (j = 0; j < my; j++)
        for (i = 0; i < mx; i++)
            for (k = 0; k < meqn; k++)
                *f++ = (float) q[((size_t) k * (my + 2 * mbc) + j + mbc) * (mx + 2 * mbc) + i + mbc];
}

/* values converted on the device, handed out patch by patch */
static const float *vtk_device_values;

static void vtk_device_value_cb(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                                int patchno, char *a)
{
    const fclaw2d_clawpatch_options_t *o = &glob->clawpatch_opt;
    const size_t per = (size_t) o->mx * o->my * o->meqn;
    (void) blockno; (void) patchno;
    memcpy(a, vtk_device_values + per * (size_t) patch->slot, per * sizeof(float));
}

typedef enum {
    VTK_POSITION, VTK_CONNECTIVITY, VTK_OFFSETS, VTK_TYPES, VTK_MPIRANK, VTK_BLOCKNO,
    VTK_PATCHNO, VTK_MEQN, VTK_NFIELDS
} vtk_field_t;

int fclaw2d_vtk_write_file(fclaw2d_global_t *glob, const char *basename, int mx, int my,
                           int meqn, double vtkspace, int vtkwrite,
                           fclaw2d_vtk_patch_data_t coordinate_cb,
                           fclaw2d_vtk_patch_data_t value_cb)
{
    fclaw2d_domain_t *d = glob->domain;
    const long long np = d->local_num_patches;
    const long long ppp = (long long) (mx + 1) * (my + 1), cpp = (long long) mx * my;
    const long long npoints = ppp * np, ncells = cpp * np, nconn = 4 * (ncells + 1);
    const int fits32 = npoints <= INT32_MAX && nconn <= INT32_MAX;
    const int isz = fits32 ? 4 : 8, ndsize = 8;
    const char *itype = fits32 ? "Int32" : "Int64";
    long long psize[VTK_NFIELDS], offset[VTK_NFIELDS + 1];
    char fname[4096], *buf;
    FILE *f;
    int field, b, p, ok;
    (void) vtkspace; (void) vtkwrite;
    psize[VTK_POSITION] = ppp * 3 * 8;
    psize[VTK_CONNECTIVITY] = cpp * 4 * isz;
    psize[VTK_OFFSETS] = cpp * isz;
    psize[VTK_TYPES] = cpp;
    psize[VTK_MPIRANK] = cpp * 4;
    psize[VTK_BLOCKNO] = cpp * 4;
    psize[VTK_PATCHNO] = cpp * isz;
    psize[VTK_MEQN] = cpp * meqn * 4;
    offset[0] = 0;
    for (field = 0; field < VTK_NFIELDS; field++)
        offset[field + 1] = offset[field] + ndsize + psize[field] * np;
    snprintf(fname, sizeof fname, "%s.vtu", basename);
    f = fopen(fname, "wb");
    if (!f) return -1;
    ok = fprintf(f, "<?xml version=\"1.0\"?>\n"
                    "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\" "
                    "header_type=\"UInt64\" >\n"
                    " <UnstructuredGrid>\n"
                    "  <Piece NumberOfPoints=\"%lld\" NumberOfCells=\"%lld\">\n"
                    "   <Points>\n"
                    "    <DataArray type=\"Float64\" Name=\"Position\" NumberOfComponents=\"3\" "
                    "format=\"appended\" offset=\"%lld\">\n"
                    "    </DataArray>\n"
                    "   </Points>\n"
                    "   <Cells>\n", npoints, ncells, offset[VTK_POSITION]) >= 0;
    ok = ok && fprintf(f, "    <DataArray type=\"%s\" Name=\"connectivity\" format=\"appended\" "
                          "offset=\"%lld\">\n    </DataArray>\n", itype, offset[VTK_CONNECTIVITY]) >= 0;
    ok = ok && fprintf(f, "    <DataArray type=\"%s\" Name=\"offsets\" format=\"appended\" "
                          "offset=\"%lld\">\n    </DataArray>\n", itype, offset[VTK_OFFSETS]) >= 0;
    ok = ok && fprintf(f, "    <DataArray type=\"UInt8\" Name=\"types\" format=\"appended\" "
                          "offset=\"%lld\">\n    </DataArray>\n", offset[VTK_TYPES]) >= 0;
    ok = ok && fprintf(f, "   </Cells>\n   <CellData Scalars=\"mpirank,blockno,patchno\" "
                          "Fields=\"meqn\">\n") >= 0;
    ok = ok && fprintf(f, "    <DataArray type=\"Int32\" Name=\"mpirank\" format=\"appended\" "
                          "offset=\"%lld\">\n    </DataArray>\n", offset[VTK_MPIRANK]) >= 0;
    ok = ok && fprintf(f, "    <DataArray type=\"Int32\" Name=\"blockno\" format=\"appended\" "
                          "offset=\"%lld\">\n    </DataArray>\n", offset[VTK_BLOCKNO]) >= 0;
    ok = ok && fprintf(f, "    <DataArray type=\"%s\" Name=\"patchno\" format=\"appended\" "
                          "offset=\"%lld\">\n    </DataArray>\n", itype, offset[VTK_PATCHNO]) >= 0;
    if (meqn == 1)
        ok = ok && fprintf(f, "    <DataArray type=\"Float32\" Name=\"meqn\" format=\"appended\" "
                              "offset=\"%lld\">\n", offset[VTK_MEQN]) >= 0;
    else
        ok = ok && fprintf(f, "    <DataArray type=\"Float32\" Name=\"meqn\" NumberOfComponents=\"%d\" "
                              "format=\"appended\" offset=\"%lld\">\n", meqn, offset[VTK_MEQN]) >= 0;
    ok = ok && fprintf(f, "    </DataArray>\n   </CellData>\n   <PointData>\n   </PointData>\n"
                          "  </Piece>\n </UnstructuredGrid>\n <AppendedData encoding=\"raw\">\n  _") >= 0;
    buf = (char *) malloc((size_t) (psize[VTK_POSITION] > psize[VTK_CONNECTIVITY] ? psize[VTK_POSITION] : psize[VTK_CONNECTIVITY]) +
                          (size_t) psize[VTK_MEQN] + 64);
    if (!buf) { fclose(f); return -1; }
    for (field = 0; ok && field < VTK_NFIELDS; field++) {
        const unsigned long long bytes = (unsigned long long) (psize[field] * np);
        ok = fwrite(&bytes, (size_t) ndsize, 1, f) == 1;
        /* patches in global order: block, then patch within the block */
        for (b = 0; ok && b < d->num_blocks; b++)
            for (p = 0; ok && p < d->blocks[b].num_patches; p++) {
                fclaw2d_patch_t *pt = &d->blocks[b].patches[p];
                const long long gp = d->blocks[b].patchno_offset + p;
                long long c, l;
                int i, j;
                switch (field) {
                case VTK_POSITION: coordinate_cb(glob, pt, b, p, buf); break;
                case VTK_CONNECTIVITY:
                    for (j = 0, c = 0; j < my; j++)
                        for (i = 0; i < mx; i++) {
                            const long long base = ppp * gp + i + (long long) j * (mx + 1);
                            const long long v[4] = {base, base + 1, base + mx + 2, base + mx + 1};
                            for (l = 0; l < 4; l++, c++) {
                                if (fits32) ((int32_t *) buf)[c] = (int32_t) v[l];
                                else ((int64_t *) buf)[c] = v[l];
                            }
                        }
                    break;
                case VTK_OFFSETS:
                    for (c = 0; c < cpp; c++) {
                        const long long v = 4 * (cpp * gp + 1) + 4 * c;
                        if (fits32) ((int32_t *) buf)[c] = (int32_t) v;
                        else ((int64_t *) buf)[c] = v;
                    }
                    break;
                case VTK_TYPES: memset(buf, 9, (size_t) cpp); break;   /* VTK_QUAD */
                case VTK_MPIRANK: for (c = 0; c < cpp; c++) ((int32_t *) buf)[c] = 0; break;
                case VTK_BLOCKNO: for (c = 0; c < cpp; c++) ((int32_t *) buf)[c] = b; break;
                case VTK_PATCHNO:
                    for (c = 0; c < cpp; c++) {
                        if (fits32) ((int32_t *) buf)[c] = (int32_t) gp;
                        else ((int64_t *) buf)[c] = gp;
                    }
                    break;
                default: value_cb(glob, pt, b, p, buf); break;
                }
                ok = fwrite(buf, (size_t) psize[field], 1, f) == 1;
            }
    }
    free(buf);
    ok = ok && fprintf(f, "\n </AppendedData>\n</VTKFile>\n") >= 0;
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : -1;
}

void fclaw2d_clawpatch_output_vtk(fclaw2d_global_t *glob, int iframe)
{
    const fclaw2d_clawpatch_options_t *o = &glob->clawpatch_opt;
    const int n = glob->domain->local_num_patches;
    char basename[4096];
    float *vals = (float *) malloc(sizeof(float) * (size_t) o->mx * o->my * o->meqn * (size_t) (n > 0 ? n : 1));
    if (!vals) {
        fprintf(stderr, "fclaw2d_clawpatch_output_vtk: out of memory\n");
        exit(1);
    }
    device_check(hipclaw_vtk_values(glob->state, 0, n, vals), "hipclaw_vtk_values");
    snprintf(basename, sizeof basename, "%s_frame_%04d", glob->prefix ? glob->prefix : "fort", iframe);
    vtk_device_values = vals;
    if (fclaw2d_vtk_write_file(glob, basename, o->mx, o->my, o->meqn, 0.0, 0,
                               fclaw2d_output_vtk_coordinate_cb, vtk_device_value_cb))
        fprintf(stderr, "fclaw2d_clawpatch_output_vtk: cannot write %s.vtu\n", basename);
    vtk_device_values = NULL;
    free(vals);
}
