/*
 * Test-only C harness: the reference's Clawpack-4.6 patch step
 * (clawpack46_step2_wrap.f and what it calls) with the reference's acoustics
 * Riemann solvers (applications/clawpack/acoustics/2d/rp/
 * clawpack46_rpn2_acoustics.f, clawpack46_rpt2_acoustics.f), compiled in place
 * by oracle/Makefile into oracle/_ref/libref46_acoustics.so.  It sets
 * common /cparam/ rho, bulk, cc, zz as the acoustics setprob does
 * (applications/clawpack/acoustics/2d/radial/setprob.f:14-20).
 *
 * Used only to pin the oracle's C restatement; never linked into the product.
 */
#include <math.h>
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

extern struct { double rho, bulk, cc, zz; } cparam_;

/* One step of one patch (meqn 3, mwaves 2, maux 0); returns ierror. */
int ref46_acoustics_step2(int mx, int my, int mbc, const int *method_in, const int *mthlim_in,
                          double *q, double dx, double dy, double dt, double rho, double bulk,
                          double *cfl, double *fm, double *fp, double *gm, double *gp)
{
    int method[7], mthlim[10], k, maxm = mx > my ? mx : my;
    int meqn = 3, maux = 0, mwaves = 2;
    int mcapa = 0, level = 0, ierror = 0, use_fwaves = 0, bcc[4] = {0, 0, 0, 0};
    int mwork = (maxm + 2 * mbc) * (12 * meqn + (meqn + 1) * mwaves + 3 * maux + 2);
    double xlower = 0.0, ylower = 0.0, t = 0.0, auxdummy = 0.0;
    double *work = (double *) calloc((size_t) mwork, sizeof(double));
    if (!work) return -1;
    for (k = 0; k < 7; k++) method[k] = method_in[k];
    for (k = 0; k < 10; k++) mthlim[k] = k < mwaves ? mthlim_in[k] : 0;
    cparam_.rho = rho;
    cparam_.bulk = bulk;
    cparam_.cc = sqrt(bulk / rho);
    cparam_.zz = rho * cparam_.cc;
    clawpack46_step2_wrap_(&maxm, &meqn, &maux, &mbc, method, mthlim, &mcapa, &mwaves,
                           &mx, &my, q, &auxdummy, &dx, &dy, &dt, cfl,
                           work, &mwork, &xlower, &ylower, &level, &t, fp, fm, gp, gm,
                           clawpack46_rpn2_, clawpack46_rpt2_, (fproc_t) 0, (fproc_t) 0,
                           clawpack46_flux2_, bcc, &ierror, &use_fwaves);
    free(work);
    return ierror;
}
