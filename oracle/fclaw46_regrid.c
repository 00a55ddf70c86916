/*
 * Oracle (test infrastructure only): plain-C restatement of the reference's
 * regrid arithmetic -- the gradient refinement criterion and the transfer of
 * the solution to new patches.  Never linked into the product.
 *
 *   tag4refinement  : src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_tag4refinement.f:1-70
 *                     with fclaw2d_clawpatch_gradient_exceeds_th
 *                     (src/patches/clawpatch/fort_user/fclaw2d_clawpatch_gradient_exceeds_th.f90,
 *                     unmapped branch) and clawpatch_tag4refinement
 *                     (src/patches/clawpatch/fclaw2d_clawpatch.cpp:800-850)
 *   tag4coarsening  : src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_tag4coarsening.f
 *   interpolate2fine: fclaw2d_clawpatch46_interpolate.f (fclaw2d_clawpatch46_fort_interpolate2fine,
 *                     manifold = 0) with the slopes of fclaw2d_clawpatch_utils.f (mth = 5)
 *   average2coarse  : fclaw2d_clawpatch46_average.f (fclaw2d_clawpatch46_fort_average2coarse,
 *                     manifold = 0)
 *   repopulate      : cb_fclaw2d_regrid_repopulate (src/fclaw2d_regrid.c:95-230): a new
 *                     patch is a copy of the same old patch, the interpolation of its
 *                     old parent, or the average of its four old children.
 *
 * The old and new leaf sets are inputs: how the forest adapts (p4est's
 * refine / coarsen / balance) is not restated here.
 */
#include "fclaw46_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define Q(a, i, j, m) (a)[((size_t) (m) * (my + 2 * mbc) + (size_t) ((j) + mbc - 1)) * (mx + 2 * mbc) + (i) + mbc - 1]

/* fclaw2d_clawpatch_gradient_exceeds_th for one interior cell, unmapped:
 * |(dq/dx, dq/dy, 0)| > threshold with centred differences. */
static int gradient_exceeds(int mx, int my, int mbc, const double *q, int mq, int i, int j,
                            double dx, double dy, double threshold)
{
    const double dx2 = 2 * dx, dy2 = 2 * dy;
    const double dqx = (Q(q, i + 1, j, mq) - Q(q, i - 1, j, mq)) / dx2;
    const double dqy = (Q(q, i, j + 1, mq) - Q(q, i, j - 1, mq)) / dy2;
    const double g3 = 0.0;
    const double ds = sqrt(dqx * dqx + dqy * dqy + g3 * g3);
    return ds > threshold;
}

int oracle_patch_exceeds(int mx, int my, int mbc, const double *q, int ivar, double dx,
                         double dy, double threshold)
{
    int i, j;
    if (threshold < 0) return 1;     /* clawpatch_tag4refinement: always refine */
    for (j = 1; j <= my; j++)
        for (i = 1; i <= mx; i++)
            if (gradient_exceeds(mx, my, mbc, q, ivar, i, j, dx, dy, threshold)) return 1;
    return 0;
}

/* interpolate2fine into fine child igrid (0..3, z-order) of a coarse patch,
 * over the loop ranges of the reference (which include one layer of the
 * fine patch's ghost cells on the side facing its siblings). */
void oracle_interpolate2fine(int mx, int my, int mbc, int meqn, const double *qcoarse,
                             double *qfine, int igrid)
{
    const int ig = igrid % 2, jg = (igrid - ig) / 2;
    const int i1 = 1 - ig, i2 = mx / 2 + (1 - ig), ic_add = ig * mx / 2;
    const int j1 = 1 - jg, j2 = my / 2 + (1 - jg), jc_add = jg * my / 2;
    int mq, i, j, ii, jj;
    for (mq = 0; mq < meqn; mq++)
        for (j = j1; j <= j2; j++)
            for (i = i1; i <= i2; i++) {
                const int ic = i + ic_add, jc = j + jc_add;
                const double qc = Q(qcoarse, ic, jc, mq);
                const double gradx = oracle_compute_slopes(qc - Q(qcoarse, ic - 1, jc, mq), Q(qcoarse, ic + 1, jc, mq) - qc);
                const double grady = oracle_compute_slopes(qc - Q(qcoarse, ic, jc - 1, mq), Q(qcoarse, ic, jc + 1, mq) - qc);
                for (ii = 1; ii <= 2; ii++)
                    for (jj = 1; jj <= 2; jj++) {
                        const double shiftx = (ii - 2 / 2.0 - 0.5) / 2;
                        const double shifty = (jj - 2 / 2.0 - 0.5) / 2;
                        Q(qfine, (i - 1) * 2 + ii, (j - 1) * 2 + jj, mq) = qc + shiftx * gradx + shifty * grady;
                    }
            }
}

void oracle_average2coarse(int mx, int my, int mbc, int meqn, double *qcoarse,
                           const double *qfine, int igrid)
{
    const int ig = igrid % 2, jg = (igrid - ig) / 2;
    const int ic_add = ig * mx / 2, jc_add = jg * mx / 2;  /* mx, as the reference writes it */
    int mq, i, j;
    for (mq = 0; mq < meqn; mq++)
        for (j = 1; j <= my / 2; j++)
            for (i = 1; i <= mx / 2; i++) {
                double sum = 0;
                sum = sum + Q(qfine, 2 * i - 1, 2 * j - 1, mq);
                sum = sum + Q(qfine, 2 * i, 2 * j - 1, mq);
                sum = sum + Q(qfine, 2 * i - 1, 2 * j, mq);
                sum = sum + Q(qfine, 2 * i, 2 * j, mq);
                Q(qcoarse, i + ic_add, j + jc_add, mq) = sum / 4;
            }
}

/* Repopulate the new leaves from the old solution.  Leaves are (level, gx,
 * gy) over the brick; a new leaf must be an old leaf, a child of one, or the
 * parent of four.  Returns 0, or -1 if some new leaf is none of these. */
int oracle_regrid_remap(int mx, int my, int mbc, int meqn, int nold, const int *lold,
                        const int *xold, const int *yold, const double *qold, int nnew,
                        const int *lnew, const int *xnew, const int *ynew, double *qnew)
{
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    int p, k, g;
    for (p = 0; p < nnew; p++) {
        int found = 0;
        double *qn = qnew + (size_t) p * psize;
        for (k = 0; k < nold && !found; k++) {
            if (lold[k] == lnew[p] && xold[k] == xnew[p] && yold[k] == ynew[p]) {
                memcpy(qn, qold + (size_t) k * psize, psize * sizeof(double));
                found = 1;
            } else if (lold[k] == lnew[p] - 1 && xold[k] == xnew[p] / 2 && yold[k] == ynew[p] / 2) {
                oracle_interpolate2fine(mx, my, mbc, meqn, qold + (size_t) k * psize, qn,
                                        (xnew[p] & 1) + 2 * (ynew[p] & 1));
                found = 1;
            }
        }
        if (found) continue;
        for (g = 0; g < 4; g++) {
            const int cx = 2 * xnew[p] + (g & 1), cy = 2 * ynew[p] + (g >> 1);
            int hit = 0;
            for (k = 0; k < nold; k++)
                if (lold[k] == lnew[p] + 1 && xold[k] == cx && yold[k] == cy) {
                    oracle_average2coarse(mx, my, mbc, meqn, qn, qold + (size_t) k * psize, g);
                    hit = 1;
                    break;
                }
            if (!hit) return -1;
        }
    }
    return 0;
}
