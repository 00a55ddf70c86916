/*
 * CPU oracle for the batched patch-update path: a plain-C restatement of the
 * reference's Clawpack-4.6 patch step, Riemann solvers, clawpatch46 ghost-fill
 * routines and their single-process orchestration.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py use it, and only as the checker or the timed CPU baseline.
 *
 * Every routine cites the reference file it restates (paths relative to the
 * reference root).  Index conventions follow the Fortran: a patch array
 * q(1-mbc:mx+mbc, 1-mbc:my+mbc, meqn) is stored with i fastest, then j, then m.
 */
#ifndef FCLAW46_ORACLE_H
#define FCLAW46_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Riemann solver families restated from the reference applications. */
enum {
    ORACLE_RP_ADVECTION = 0, /* applications/clawpack/advection/2d/all/clawpack46_rpn2adv.f, _rpt2adv.f */
    ORACLE_RP_SHALLOW   = 1, /* applications/clawpack/shallow/2d/rp/clawpack46_rpn2.f, _rpt2.f (efix on) */
    ORACLE_RP_EULER     = 2, /* applications/clawpack/euler/2d/rp/clawpack46_rpn2_euler4.f, _rpt2_euler4.f */
    ORACLE_RP_ACOUSTICS = 3, /* applications/clawpack/acoustics/2d/rp/clawpack46_rpn2_acoustics.f, _rpt2_acoustics.f */
    ORACLE_RP_TRACERS   = 4  /* meqn = mwaves independent colour equations, tracer k advected with
                                velocity c_k (u, v), c_k = (meqn - k) / meqn, k = 0..meqn-1: the
                                advection solvers above applied per component (a test solver with
                                up to 10 wave families; no reference counterpart) */
};

typedef struct oracle_claw {
    int solver;       /* ORACLE_RP_* */
    int meqn, maux, mwaves;
    int method[7];    /* Clawpack method array; method(2)=order, method(3)=transverse */
    int mthlim[10];   /* limiter per wave family (0..6) */
    double grav;      /* shallow water: common /cparam/ grav */
    double gamma;     /* Euler: common /cparam/ gamma, gamma1 = gamma - 1 */
    double p[8];      /* acoustics: common /cparam/ rho, bulk, cc, zz */
} oracle_claw_t;

/* One patch step: clawpack46_step2_wrap.f (step2 + flux2 + q update).
 * q is updated in place (interior and ghost cells 1-mbc..mx+mbc-1, as the
 * reference does).  fm/fp/gm/gp may be NULL; if given they receive the
 * reference's flux arrays (size meqn*(mx+2mbc)*(my+2mbc)).
 * Returns 0 on success; *cfl receives cflgrid. */
int oracle_step2_patch(const oracle_claw_t *claw, int mx, int my, int mbc,
                       double *q, const double *aux,
                       double dx, double dy, double dt, double *cfl,
                       double *fm, double *fp, double *gm, double *gp);

/* Limiter only (clawpack46_inlinelimiter.f) on a 1D slice of waves laid out as
 * wave(1-mbc:maxm+mbc, meqn, mwaves) and s(1-mbc:maxm+mbc, mwaves). */
void oracle_inlinelimiter(int maxm, int meqn, int mwaves, int mbc, int mx,
                          double *wave, const double *s, const int *mthlim);

/* One-slice Riemann solves (rpn2 over i=2-mbc..mx+mbc) for unit checks. */
void oracle_rpn2(const oracle_claw_t *claw, int ixy, int maxm, int mbc, int mx,
                 const double *q1d, const double *aux1d,
                 double *wave, double *s, double *amdq, double *apdq);

/* ---------------- forest-level oracle (single process) ---------------- */

typedef struct oracle_domain {
    int npatch;
    const int *level;   /* [npatch] refinement level */
    const int *gx;      /* [npatch] patch column at its level, over the whole brick */
    const int *gy;      /* [npatch] patch row at its level, over the whole brick */
    int nbx, nby;       /* brick of nbx x nby unit blocks */
    int periodic_x, periodic_y;
    int mx, my, mbc, meqn, maux;
    int mthbc[4];       /* boundary type per domain face: 1 extrapolation, 3 wall */
    double ax, bx, ay, by;
} oracle_domain_t;

/* Ghost update over all levels present: fclaw2d_ghost_update_async
 * (src/fclaw2d_ghost_fill.c:611-830) with one process. q: [npatch][patch]. */
int oracle_ghost_update(const oracle_domain_t *d, double *q);

/* Update every patch with oracle_step2_patch at a common dt; returns max cfl. */
double oracle_update_all(const oracle_domain_t *d, const oracle_claw_t *claw,
                         double *q, const double *aux, double dt);

/* nsteps of (update all levels, ghost update) at fixed dt (subcycle = F). */
double oracle_advance(const oracle_domain_t *d, const oracle_claw_t *claw,
                      double *q, const double *aux, double dt, int nsteps);

/* Same with the reference's time synchronisation switched on or off. */
double oracle_advance_ts(const oracle_domain_t *d, const oracle_claw_t *claw,
                         double *q, const double *aux, double dt, int nsteps, int time_sync);

/* One fclaw2d_advance_all_levels(t, dt) with subcycle = T: the coarsest level
 * steps once with dt, level L 2^(L-minlevel) times with dt/2^(L-minlevel),
 * with time-interpolated ghost exchanges and (if time_sync) the
 * synchronisations in between.  Returns the maximum CFL number. */
double oracle_advance_subcycle(const oracle_domain_t *d, const oracle_claw_t *claw, double *q,
                               const double *aux, double dt, int time_sync);

/* ---------------- time synchronisation (conservation fix-up) ----------------
 * Flux registers of one patch (mx == my): reg[((face*4 + kind)*meqn + m)*mx + pos]
 * with face 0..3 (left, right, bottom, top), pos = j-1 (faces 0/1) or i-1
 * (faces 2/3).  Kinds: the step's fm and fp at that edge (gm/gp on faces 2/3)
 * and the conservative flux at the edge's interior and exterior cell. */
enum { ORACLE_REG_FM = 0, ORACLE_REG_FP = 1, ORACLE_REG_EFIN = 2, ORACLE_REG_EFOUT = 3 };
size_t oracle_sync_register_count(int mx, int meqn);
void oracle_cons_flux(const oracle_claw_t *claw, int idir, const double *q, const double *auxc,
                      double *flux);
void oracle_store_flux(const oracle_claw_t *claw, int mx, int my, int mbc, const double *q,
                       const double *aux, double dt, double dx, double dy, double *reg);
void oracle_accumulate_waves(int mx, int my, int mbc, int meqn, double dt, double dx, double dy,
                             const double *fp, const double *fm, const double *gp,
                             const double *gm, double *reg);
void oracle_time_sync_f2c(int mx, int my, int mbc, int meqn, double *qcoarse,
                          const double *rcoarse, const double *rfine, int iface, int ti, int tj,
                          double dx, double dy);
void oracle_time_sync_samesize(int mx, int my, int mbc, int meqn, double *qthis,
                               const double *rthis, const double *rnbr, int iface, double dx,
                               double dy);
/* fclaw2d_time_sync over all levels; registers are left zero. */
int oracle_time_sync(const oracle_domain_t *d, double *q, double *reg);

/* ---------------- regrid ---------------- */
/* fclaw2d_clawpatch_compute_slopes, mth = 5 (fclaw2d_clawpatch_utils.f:45-63) */
double oracle_compute_slopes(double sl, double sr);
/* 1 if some interior cell's |grad q(ivar)| exceeds threshold (always 1 for a
 * negative threshold): the gradient criterion of tag4refinement/coarsening. */
int oracle_patch_exceeds(int mx, int my, int mbc, const double *q, int ivar, double dx,
                         double dy, double threshold);
void oracle_interpolate2fine(int mx, int my, int mbc, int meqn, const double *qcoarse,
                             double *qfine, int igrid);
void oracle_average2coarse(int mx, int my, int mbc, int meqn, double *qcoarse,
                           const double *qfine, int igrid);
int oracle_regrid_remap(int mx, int my, int mbc, int meqn, int nold, const int *lold,
                        const int *xold, const int *yold, const double *qold, int nnew,
                        const int *lnew, const int *xnew, const int *ynew, double *qnew);

/* ---------------- ghost-patch message ---------------- */
size_t oracle_ghost_pack_elems(int mx, int my, int mbc, int meqn, int refratio, int packarea,
                               int packregisters);
long oracle_local_ghost_pack(int mx, int my, int mbc, int meqn, int mint, double *q,
                             double *area, double *pack, int packmode);
long oracle_register_pack(int mx, int my, int meqn, double dx, double dy, double *reg,
                          double *pack, int unpack);

/* Kahan mass sum per field in patch order:
 * src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_diagnostics.f:13-55 */
void oracle_mass(const oracle_domain_t *d, const double *q, double *sum);

/* Individual ghost-fill kernels on two patch arrays (translation transforms).
 * For same-size: neighbor index = this index + (ti, tj).
 * For half-size (this = coarse): first fine child of coarse cell (ic,jc) is
 * (2*ic - 1 + ti, 2*jc - 1 + tj) in the fine patch. */
void oracle_copy_face(int mx, int my, int mbc, int meqn, double *qthis,
                      const double *qnbr, int iface, int ti, int tj);
void oracle_copy_corner(int mx, int my, int mbc, int meqn, double *qthis,
                        const double *qnbr, int icorner, int ti, int tj);
void oracle_average_face(int mx, int my, int mbc, int meqn, double *qcoarse,
                         const double *qfine, int idir, int iface_coarse,
                         int ti, int tj);
void oracle_average_corner(int mx, int my, int mbc, int meqn, double *qcoarse,
                           const double *qfine, int icorner, int ti, int tj);
void oracle_interpolate_face(int mx, int my, int mbc, int meqn,
                             const double *qcoarse, double *qfine, int idir,
                             int iface_coarse, int ti, int tj);
/* fclaw2d_clawpatch46_fort_timeinterp on one patch (the interior band of width mbc) */
void oracle_timeinterp_patch(int mx, int my, int mbc, int meqn, const double *qcurr,
                             const double *qlast, double *qinterp, double alpha);
void oracle_interpolate_corner(int mx, int my, int mbc, int meqn,
                               const double *qcoarse, double *qfine,
                               int icorner, int ti, int tj);
void oracle_bc2(int mx, int my, int mbc, int meqn, double *q,
                const int *mthbc_patch);

#ifdef __cplusplus
}
#endif
#endif
