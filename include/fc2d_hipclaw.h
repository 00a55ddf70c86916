/*
 * fc2d_hipclaw: the MI355X solver plugin.  Drop-in for the reference's
 * fc2d_cudaclaw plugin (src/solvers/fc2d_cudaclaw/fc2d_cudaclaw.h,
 * fc2d_cudaclaw_cuda.h, fc2d_cudaclaw_options.h): same vtable slots, options
 * fields, buffering of patches in single_step_update and return values.
 *
 * Differences a maintainer switching from fc2d_cudaclaw must know:
 *  - cuda_rpn2 / cuda_rpt2 device function pointers are replaced by
 *    `riemann_solver`, one of the HIPCLAW_RP_* functors compiled into the
 *    backend, or `rp_solver`, the descriptor of the application's own functor
 *    compiled into the step kernel (hipclaw_solver.h; FP32 and FP64
 *    instantiations);
 *  - patch data stay on the GPU between steps; host views are synchronised
 *    with fc2d_hipclaw_download / fc2d_hipclaw_upload;
 *  - options add `precision` (HIPCLAW_FP64 or HIPCLAW_FP32_BULK) and `tile`.
 */
#ifndef FC2D_HIPCLAW_H
#define FC2D_HIPCLAW_H

#include "fclaw2d_hip_core.h"
#include "hipclaw.h"
#include "hipclaw_solver.h"

#ifdef __cplusplus
extern "C" {
#endif

#define FC2D_HIPCLAW_MWAVES 10   /* as FC2D_CUDACLAW_MWAVES */

/* ---- options (src/solvers/fc2d_cudaclaw/fc2d_cudaclaw_options.h:42-67) ---- */
typedef struct fc2d_hipclaw_options {
    int mwaves;
    int order[2];
    int mthlim[FC2D_HIPCLAW_MWAVES];
    int mthbc[4];
    int method[7];
    int mcapa;
    int src_term;
    int use_fwaves;
    int ascii_out;
    int vtk_out;
    int buffer_len;        /* patches per device batch (reference default 1) */
    int precision;         /* HIPCLAW_FP64 or HIPCLAW_FP32_BULK */
    int tile;              /* 0 = tuned default */
    double grav, gamma;    /* physical constants of the compiled Riemann solvers */
    double rp_params[8];   /* constants of an application's own solver (vtable rp_solver) */
    int is_registered;
} fc2d_hipclaw_options_t;

/* Defaults of fc2d_cudaclaw_options.c:36-71: order 2 2, mthlim 1, mthbc 1,
 * mcapa 0, src_term 0, use_fwaves 0, buffer_len 1. */
void fc2d_hipclaw_options_default(fc2d_hipclaw_options_t *opt, int mwaves);
void fc2d_hipclaw_options_store(fclaw2d_global_t *glob, fc2d_hipclaw_options_t *opt);
fc2d_hipclaw_options_t *fc2d_hipclaw_get_options(fclaw2d_global_t *glob);
/* fc2d_cudaclaw_check: returns 0 (FCLAW_NOEXIT) or 1 (FCLAW_EXIT_ERROR) */
int fc2d_hipclaw_check(const fc2d_hipclaw_options_t *opt);
/* cudaclaw_check_parameters (fc2d_cudaclaw_options.c:110-115) */
int hipclaw_check_parameters(int mwaves);
/* cudaclaw_set_method_parameters: order, limiters and f-wave flag for later
 * batches of the current global */
void hipclaw_set_method_parameters(fclaw2d_global_t *glob, const int order[], const int mthlim[],
                                   int mwaves, int use_fwaves);

/* ---- vtable (src/solvers/fc2d_cudaclaw/fc2d_cudaclaw.h:141-165) ---- */
typedef void (*hipclaw_src2_t)(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                               int patchno, double t, double dt);
typedef void (*hipclaw_b4step2_t)(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                                  int patchno, double t, double dt);
typedef void (*hipclaw_fort_qinit_t)(const int *maxmx, const int *maxmy, const int *meqn,
                                     const int *mbc, const int *mx, const int *my,
                                     const double *xlower, const double *ylower,
                                     const double *dx, const double *dy, double q[],
                                     const int *maux, double aux[]);
typedef void (*hipclaw_fort_setaux_t)(const int *maxmx, const int *maxmy, const int *mbc,
                                      const int *mx, const int *my, const double *xlower,
                                      const double *ylower, const double *dx, const double *dy,
                                      const int *maux, double aux[]);
typedef void (*hipclaw_fort_bc2_t)(const int *maxmx, const int *maxmy, const int *meqn,
                                   const int *mbc, const int *mx, const int *my,
                                   const double *xlower, const double *ylower, const double *dx,
                                   const double *dy, double q[], const int *maux,
                                   const double aux[], const double *t, const double *dt,
                                   const int mthbc[]);
typedef void (*hipclaw_fort_src2_t)(const int *maxmx, const int *maxmy, const int *meqn,
                                    const int *mbc, const int *mx, const int *my,
                                    const double *xlower, const double *ylower,
                                    const double *dx, const double *dy, double q[],
                                    const int *maux, double aux[], const double *t,
                                    const double *dt);
typedef void (*hipclaw_fort_setprob_t)(void);

typedef struct fc2d_hipclaw_vtable {
    hipclaw_b4step2_t b4step2;       /* not called on the GPU path (as cudaclaw) */
    hipclaw_src2_t src2;
    hipclaw_fort_setprob_t fort_setprob;
    hipclaw_fort_bc2_t fort_bc2;     /* host-side BC for host views; the device uses mthbc */
    hipclaw_fort_qinit_t fort_qinit;
    hipclaw_fort_setaux_t fort_setaux;
    hipclaw_fort_src2_t fort_src2;
    int riemann_solver;              /* HIPCLAW_RP_*: a built-in solver ... */
    const hipclaw_solver_t *rp_solver; /* ... or the application's own (hipclaw_solver.h),
                                          taking precedence: replaces cuda_rpn2 / cuda_rpt2 */
    int is_set;
} fc2d_hipclaw_vtable_t;

/* fc2d_cudaclaw_solver_initialize (fc2d_cudaclaw.cpp:405-465): installs
 * single_step_update, initialize, setup and physical_bc in the patch vtable. */
void fc2d_hipclaw_solver_initialize(fclaw2d_global_t *glob);
fc2d_hipclaw_vtable_t *fc2d_hipclaw_vt(fclaw2d_global_t *glob);

/* fc2d_cudaclaw_initialize_GPUs (cudaclaw_initialize.cu:37-74) */
void fc2d_hipclaw_initialize_GPUs(fclaw2d_global_t *glob);
/* fc2d_cudaclaw_allocate_buffers / deallocate_buffers (cudaclaw_allocate.cu:112-157):
 * creates the device state for the current domain (all patches) */
void fc2d_hipclaw_allocate_buffers(fclaw2d_global_t *glob);
void fc2d_hipclaw_deallocate_buffers(fclaw2d_global_t *glob);

/* Host-view synchronisation (no counterpart in the reference, whose patch data
 * always live on the host). */
void fc2d_hipclaw_upload(fclaw2d_global_t *glob);
void fc2d_hipclaw_download(fclaw2d_global_t *glob);

/* User access to solver functions (fc2d_cudaclaw.h:177-229), on host views. */
void fc2d_hipclaw_setprob(fclaw2d_global_t *glob);
void fc2d_hipclaw_setaux(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                         int patchno);
void fc2d_hipclaw_qinit(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                        int patchno);
void fc2d_hipclaw_b4step2(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                          int patchno, double t, double dt);
void fc2d_hipclaw_bc2(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno, int patchno,
                      double t, double dt, int intersects_bc[], int time_interp);
void fc2d_hipclaw_src2(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int blockno,
                       int patchno, double t, double dt);

/* ---- batching (fc2d_cudaclaw_cuda.h:78-110) ---- */
/* cudaclaw_store_buffer: record patch `patchno`'s slot in batch position iter % buffer_len */
void hipclaw_store_buffer(fclaw2d_global_t *glob, fclaw2d_patch_t *patch, int patchno,
                          int count, int iter, int *slot_buffer);
/* cudaclaw_step2_batch: one device step of `batch_size` buffered patches;
 * returns their maximum Courant number */
double hipclaw_step2_batch(fclaw2d_global_t *glob, const int *slot_buffer, int batch_size,
                           double t, double dt);

#ifdef __cplusplus
}
#endif
#endif
