/*
 * Test-only C harness that drives the reference's own Clawpack-4.6 Fortran
 * (compiled in place from the reference tree by oracle/Makefile into
 * oracle/_ref/libref46.so).  It plays the part of the reference's
 * clawpack46_step2 / clawpack46_bc2 C wrappers
 * (src/solvers/fc2d_clawpack4.6/fc2d_clawpack46.cpp:95-145, 259-400): it sets
 * the Riemann solvers' common /cparam/, sizes the work array and passes the
 * rpn2/rpt2/flux2 procedure arguments.
 *
 * Used only to pin the oracle's C restatement; never linked into the product.
 */
#include <stdlib.h>

typedef void (*fproc_t)(void);

extern void clawpack46_step2_wrap_(int *maxm, int *meqn, int *maux, int *mbc,
                                   int *method, int *mthlim, int *mcapa, int *mwaves,
                                   int *mx, int *my, double *qold, double *aux,
                                   double *dx, double *dy, double *dt, double *cfl,
                                   double *work, int *mwork, double *xlower, double *ylower,
                                   int *level, double *t, double *fp, double *fm,
                                   double *gp, double *gm, fproc_t rpn2, fproc_t rpt2,
                                   fproc_t rpn2fw, fproc_t rpt2fw, fproc_t flux2,
                                   int *block_corner_count, int *ierror, int *use_fwaves);
extern void clawpack46_flux2_(void);
extern void clawpack46_rpn2_(void);
extern void clawpack46_rpt2_(void);
extern void clawpack46_rpn2_euler4_(void);
extern void clawpack46_rpt2_euler4_(void);
extern void clawpack46_rpn2adv_(void);
extern void clawpack46_rpt2adv_(void);
extern void clawpack46_bc2_default_(int *maxmx, int *maxmy, int *meqn, int *mbc,
                                    int *mx, int *my, double *xlower, double *ylower,
                                    double *dx, double *dy, double *q, int *maux,
                                    double *aux, double *t, double *dt, int *mthbc);
extern void clawpack46_inlinelimiter_(int *maxm, int *meqn, int *mwaves, int *mbc,
                                      int *mx, double *wave, double *s, int *mthlim);

/* common /cparam/: grav (shallow) or gamma, gamma1 (Euler) */
extern struct { double v[2]; } cparam_;

/* solver: 0 advection, 1 shallow water, 2 Euler (same numbering as the oracle) */
int ref46_step2(int solver, int mx, int my, int mbc, int meqn, int maux, int mwaves,
                const int *method_in, const int *mthlim_in, double *q, double *aux,
                double dx, double dy, double dt, double grav, double gamma,
                double *cfl, double *fm, double *fp, double *gm, double *gp)
{
    int method[7], mthlim[10], k, maxm = mx > my ? mx : my;
    int mcapa = 0, level = 0, ierror = 0, use_fwaves = 0, bcc[4] = {0, 0, 0, 0};
    int mwork = (maxm + 2 * mbc) * (12 * meqn + (meqn + 1) * mwaves + 3 * maux + 2);
    double xlower = 0.0, ylower = 0.0, t = 0.0, auxdummy = 0.0;
    double *work = (double *) calloc((size_t) mwork, sizeof(double));
    fproc_t rpn2, rpt2;
    if (!work) return -1;
    for (k = 0; k < 7; k++) method[k] = method_in[k];
    for (k = 0; k < 10; k++) mthlim[k] = k < mwaves ? mthlim_in[k] : 0;
    if (solver == 0) { rpn2 = clawpack46_rpn2adv_; rpt2 = clawpack46_rpt2adv_; }
    else if (solver == 1) { rpn2 = clawpack46_rpn2_; rpt2 = clawpack46_rpt2_; cparam_.v[0] = grav; }
    else { rpn2 = clawpack46_rpn2_euler4_; rpt2 = clawpack46_rpt2_euler4_;
           cparam_.v[0] = gamma; cparam_.v[1] = gamma - 1.0; }
    clawpack46_step2_wrap_(&maxm, &meqn, &maux, &mbc, method, mthlim, &mcapa, &mwaves,
                           &mx, &my, q, maux > 0 ? aux : &auxdummy, &dx, &dy, &dt, cfl,
                           work, &mwork, &xlower, &ylower, &level, &t, fp, fm, gp, gm,
                           rpn2, rpt2, (fproc_t) 0, (fproc_t) 0, clawpack46_flux2_,
                           bcc, &ierror, &use_fwaves);
    free(work);
    return ierror;
}

void ref46_bc2(int mx, int my, int mbc, int meqn, double *q, const int *mthbc_in)
{
    int mthbc[4] = {mthbc_in[0], mthbc_in[1], mthbc_in[2], mthbc_in[3]};
    int maux = 0;
    double xlower = 0, ylower = 0, dx = 1, dy = 1, t = 0, dt = 0, aux = 0;
    clawpack46_bc2_default_(&mx, &my, &meqn, &mbc, &mx, &my, &xlower, &ylower, &dx, &dy,
                            q, &maux, &aux, &t, &dt, mthbc);
}

void ref46_inlinelimiter(int maxm, int meqn, int mwaves, int mbc, int mx,
                         double *wave, double *s, const int *mthlim_in)
{
    int mthlim[10], k;
    for (k = 0; k < 10; k++) mthlim[k] = k < mwaves ? mthlim_in[k] : 0;
    clawpack46_inlinelimiter_(&maxm, &meqn, &mwaves, &mbc, &mx, wave, s, mthlim);
}
