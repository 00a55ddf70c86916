/*
 * Oracle (test infrastructure only): plain-C restatement of the reference's
 * time synchronisation -- the coarse/fine and same-size flux-register
 * corrections that make an AMR step conservative.  Never linked into the
 * product.
 *
 *   store_flux        : src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_time_sync.f:37-148
 *                       (called before the step, fc2d_clawpack46.cpp:317-341)
 *   accumulate_waves  : clawpack46_time_sync.f:155-214 (called after the step,
 *                       fc2d_clawpack46.cpp:376-392)
 *   f2c               : clawpack46_time_sync.f:249-508, index maps of
 *                       fclaw2d_clawpatch_transform_face_half (src/forestclaw2d.c:798-834,
 *                       same-block case) and fc46_skip_this_grid (:705-735)
 *   samesize          : clawpack46_time_sync.f:513-700
 *   registers, areas,
 *   edge lengths      : src/patches/clawpatch/fclaw2d_clawpatch_conservation.c:71-110,
 *                       src/patches/clawpatch/fclaw2d_clawpatch_utils.f:259-308 (manifold = F)
 *   conservative flux : shallow water applications/clawpack/shallow/2d/radialdam/rpn2cons_update.f:6-47;
 *                       advection applications/clawpack/advection/2d/adv_order3/user_4.6/rpn2cons_update.f:1-27
 *                       with the velocity taken from the cell's aux; Euler has no
 *                       reference routine: the physical Euler flux is used.
 *
 * Registers of one patch: reg[((face * 4 + kind) * meqn + m) * mx + pos],
 * face 0..3 (left, right, bottom, top), kind ORACLE_REG_*, pos = j-1 on faces
 * 0/1 and i-1 on faces 2/3 (mx == my).  FM/FP on faces 0/1 are the
 * reference's fm_left/fp_left, fm_right/fp_right; on faces 2/3 gm/gp bottom
 * and top.  EFIN/EFOUT are edge_fluxes(.,.,0) and (.,.,1).
 */
#include "fclaw46_oracle.h"

#include <math.h>
#include <stddef.h>

#define Q(a, i, j, m) (a)[((size_t) (m) * (my + 2 * mbc) + (size_t) ((j) + mbc - 1)) * (mx + 2 * mbc) + (i) + mbc - 1]
#define REG(r, k, kind, m, pos) (r)[(((size_t) (k) * 4 + (kind)) * meqn + (m)) * mx + (pos)]

size_t oracle_sync_register_count(int mx, int meqn) { return (size_t) 16 * mx * meqn; }

void oracle_cons_flux(const oracle_claw_t *claw, int idir, const double *q, const double *auxc,
                      double *flux)
{
    if (claw->solver == ORACLE_RP_SHALLOW) {
        /* rpn2cons_update.f:27-46 */
        const int mu = idir == 0 ? 1 : 2, mv = idir == 0 ? 2 : 1;
        const double hun = q[mu], hut = q[mv], h = q[0];
        const double un = hun / h;
        flux[0] = hun;
        flux[mu] = (hun * hun) / h + (0.5 * claw->grav) * (h * h);
        flux[mv] = un * hut;
    } else if (claw->solver == ORACLE_RP_ADVECTION) {
        const double urot = auxc[idir];
        int m;
        for (m = 0; m < claw->meqn; m++) flux[m] = urot * q[m];
    } else {
        /* Euler: (rho u_n, rho u_n^2 + p, rho u_n u_t, u_n (E + p)) */
        const int mu = idir == 0 ? 1 : 2, mv = idir == 0 ? 2 : 1;
        const double rho = q[0], un = q[mu] / rho, ut = q[mv] / rho, en = q[3];
        const double p = (claw->gamma - 1.0) * (en - 0.5 * rho * (un * un + ut * ut));
        flux[0] = q[mu];
        flux[mu] = q[mu] * un + p;
        flux[mv] = q[mu] * ut;
        flux[3] = un * (en + p);
    }
}

/* clawpack46_time_sync_store_flux: edge_fluxes += dt*el*f(q) at the first
 * interior (k=0) and the first ghost (k=1) cell of every edge. */
void oracle_store_flux(const oracle_claw_t *claw, int mx, int my, int mbc, const double *q,
                       const double *aux, double dt, double dx, double dy, double *reg)
{
    const int meqn = claw->meqn, maux = claw->maux;
    double qv[8], av[8], fl[8];
    int i, j, k, m;
    for (j = 1; j <= my; j++) {
        for (k = 0; k <= 1; k++) {      /* left: q(1-k, j) */
            for (m = 0; m < maux; m++) av[m] = Q(aux, 1 - k, j, m);
            for (m = 0; m < meqn; m++) qv[m] = Q(q, 1 - k, j, m);
            oracle_cons_flux(claw, 0, qv, av, fl);
            for (m = 0; m < meqn; m++)
                REG(reg, 0, ORACLE_REG_EFIN + k, m, j - 1) += dt * dy * fl[m];
        }
        for (k = 0; k <= 1; k++) {      /* right: q(mx+k, j) */
            for (m = 0; m < maux; m++) av[m] = Q(aux, mx + k, j, m);
            for (m = 0; m < meqn; m++) qv[m] = Q(q, mx + k, j, m);
            oracle_cons_flux(claw, 0, qv, av, fl);
            for (m = 0; m < meqn; m++)
                REG(reg, 1, ORACLE_REG_EFIN + k, m, j - 1) += dt * dy * fl[m];
        }
    }
    for (i = 1; i <= mx; i++) {
        for (k = 0; k <= 1; k++) {      /* bottom: q(i, 1-k) */
            for (m = 0; m < maux; m++) av[m] = Q(aux, i, 1 - k, m);
            for (m = 0; m < meqn; m++) qv[m] = Q(q, i, 1 - k, m);
            oracle_cons_flux(claw, 1, qv, av, fl);
            for (m = 0; m < meqn; m++)
                REG(reg, 2, ORACLE_REG_EFIN + k, m, i - 1) += dt * dx * fl[m];
        }
        for (k = 0; k <= 1; k++) {      /* top: q(i, my+k) */
            for (m = 0; m < maux; m++) av[m] = Q(aux, i, my + k, m);
            for (m = 0; m < meqn; m++) qv[m] = Q(q, i, my + k, m);
            oracle_cons_flux(claw, 1, qv, av, fl);
            for (m = 0; m < meqn; m++)
                REG(reg, 3, ORACLE_REG_EFIN + k, m, i - 1) += dt * dx * fl[m];
        }
    }
}

/* clawpack46_time_sync_accumulate_waves: the step's fp/fm (gp/gm) at the four
 * edges, scaled by dt*dy (dt*dx), fp with its sign flipped. */
void oracle_accumulate_waves(int mx, int my, int mbc, int meqn, double dt, double dx, double dy,
                             const double *fp, const double *fm, const double *gp,
                             const double *gm, double *reg)
{
    int i, j, m;
    for (j = 1; j <= my; j++)
        for (m = 0; m < meqn; m++) {
            REG(reg, 0, ORACLE_REG_FP, m, j - 1) = REG(reg, 0, ORACLE_REG_FP, m, j - 1) - dt * dy * Q(fp, 1, j, m);
            REG(reg, 0, ORACLE_REG_FM, m, j - 1) = REG(reg, 0, ORACLE_REG_FM, m, j - 1) + dt * dy * Q(fm, 1, j, m);
            REG(reg, 1, ORACLE_REG_FP, m, j - 1) = REG(reg, 1, ORACLE_REG_FP, m, j - 1) - dt * dy * Q(fp, mx + 1, j, m);
            REG(reg, 1, ORACLE_REG_FM, m, j - 1) = REG(reg, 1, ORACLE_REG_FM, m, j - 1) + dt * dy * Q(fm, mx + 1, j, m);
        }
    for (i = 1; i <= mx; i++)
        for (m = 0; m < meqn; m++) {
            REG(reg, 2, ORACLE_REG_FP, m, i - 1) = REG(reg, 2, ORACLE_REG_FP, m, i - 1) - dt * dx * Q(gp, i, 1, m);
            REG(reg, 2, ORACLE_REG_FM, m, i - 1) = REG(reg, 2, ORACLE_REG_FM, m, i - 1) + dt * dx * Q(gm, i, 1, m);
            REG(reg, 3, ORACLE_REG_FP, m, i - 1) = REG(reg, 3, ORACLE_REG_FP, m, i - 1) - dt * dx * Q(gp, i, my + 1, m);
            REG(reg, 3, ORACLE_REG_FM, m, i - 1) = REG(reg, 3, ORACLE_REG_FM, m, i - 1) + dt * dx * Q(gm, i, my + 1, m);
        }
}

/* clawpack46_fort_time_sync_f2c for one (coarse, fine) pair across coarse face
 * iface.  (ti, tj): the first fine child of coarse cell (ic, jc) is
 * (2ic-1+ti, 2jc-1+tj) in the fine patch's indices.  The fine register pair
 * covering a coarse edge cell is summed lower index first, as the reference
 * builds its dummy array; cells the fine patch does not cover are left alone
 * (fc46_skip_this_grid). */
void oracle_time_sync_f2c(int mx, int my, int mbc, int meqn, double *qc, const double *rc,
                          const double *rf, int iface, int ti, int tj, double dx, double dy)
{
    const double areac = dx * dy;
    int m, c;
    for (m = 0; m < meqn; m++)
        for (c = 1; c <= (iface < 2 ? my : mx); c++) {
            double deltac = 0.0, flu, ef;
            const int f0 = 2 * c - 1 + (iface < 2 ? tj : ti);  /* fine rows (columns) f0, f0+1 */
            const int kf = iface ^ 1;        /* the fine patch's face on this interface */
            if (f0 < 1 || f0 + 1 > (iface < 2 ? my : mx)) {
                /* not covered by this fine patch: the reference adds 0/areac */
            } else if (iface == 0 || iface == 2) {
                ef = REG(rf, kf, ORACLE_REG_EFOUT, m, f0 - 1) + REG(rf, kf, ORACLE_REG_EFOUT, m, f0);
                flu = REG(rf, kf, ORACLE_REG_FP, m, f0 - 1) + REG(rf, kf, ORACLE_REG_FP, m, f0);
                deltac = (ef - REG(rc, iface, ORACLE_REG_EFIN, m, c - 1)) -
                         (flu - REG(rc, iface, ORACLE_REG_FP, m, c - 1));
            } else {
                ef = REG(rf, kf, ORACLE_REG_EFOUT, m, f0 - 1) + REG(rf, kf, ORACLE_REG_EFOUT, m, f0);
                flu = REG(rf, kf, ORACLE_REG_FM, m, f0 - 1) + REG(rf, kf, ORACLE_REG_FM, m, f0);
                deltac = (REG(rc, iface, ORACLE_REG_EFIN, m, c - 1) - ef) -
                         (flu - REG(rc, iface, ORACLE_REG_FM, m, c - 1));
            }
            if (iface == 0) Q(qc, 1, c, m) = Q(qc, 1, c, m) + deltac / areac;
            else if (iface == 1) Q(qc, mx, c, m) = Q(qc, mx, c, m) + deltac / areac;
            else if (iface == 2) Q(qc, c, 1, m) = Q(qc, c, 1, m) + deltac / areac;
            else Q(qc, c, my, m) = Q(qc, c, my, m) + deltac / areac;
        }
}

/* clawpack46_fort_time_sync_samesize: half of the mismatch between this
 * patch's and its same-size neighbour's edge flux goes to this patch. */
void oracle_time_sync_samesize(int mx, int my, int mbc, int meqn, double *qt, const double *rt,
                               const double *rn, int iface, double dx, double dy)
{
    const double area = dx * dy;
    const int kn = iface ^ 1;
    int m, c;
    for (m = 0; m < meqn; m++)
        for (c = 1; c <= (iface < 2 ? my : mx); c++) {
            const double ef = REG(rn, kn, ORACLE_REG_EFOUT, m, c - 1);
            double delta;
            if (iface == 0 || iface == 2)
                delta = (ef - REG(rt, iface, ORACLE_REG_EFIN, m, c - 1)) -
                        (REG(rn, kn, ORACLE_REG_FP, m, c - 1) - REG(rt, iface, ORACLE_REG_FP, m, c - 1));
            else
                delta = (REG(rt, iface, ORACLE_REG_EFIN, m, c - 1) - ef) -
                        (REG(rn, kn, ORACLE_REG_FM, m, c - 1) - REG(rt, iface, ORACLE_REG_FM, m, c - 1));
            if (iface == 0) Q(qt, 1, c, m) = Q(qt, 1, c, m) + 0.5 * delta / area;
            else if (iface == 1) Q(qt, mx, c, m) = Q(qt, mx, c, m) + 0.5 * delta / area;
            else if (iface == 2) Q(qt, c, 1, m) = Q(qt, c, 1, m) + 0.5 * delta / area;
            else Q(qt, c, my, m) = Q(qt, c, my, m) + 0.5 * delta / area;
        }
}
