/*
 * Oracle (test infrastructure only): plain-C restatement of the reference's
 * ghost-patch message, the data one rank sends another for each patch that
 * borders it.  Never linked into the product.
 *
 *   message size      : clawpatch_ghost_pack_elems
 *                       (src/patches/clawpatch/fclaw2d_clawpatch.cpp:1020-1060):
 *                       (wg - hole) (meqn + packarea + packextra) + frsize, with
 *                       wg = (mx + 2 mbc)(my + 2 mbc), hole = (mx - 2 mint)(my - 2 mint),
 *                       mint = refratio mbc, frsize = 2 (4 meqn + 2)(mx + my) when
 *                       the time-sync registers travel
 *   q (and area) band : fclaw2d_clawpatch46_fort_local_ghost_pack
 *                       (src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_local_ghost_pack.f:1-140)
 *   registers         : fclaw2d_clawpatch_time_sync_pack_registers
 *                       (src/patches/clawpatch/fclaw2d_clawpatch_conservation.c:112-195),
 *                       with the non-manifold edge lengths and areas of
 *                       clawpatch_time_sync_setup (src/patches/clawpatch/fclaw2d_clawpatch_utils.f:259-308)
 *
 * Registers are in the oracle's layout reg[((face * 4 + kind) * meqn + m) * mx + pos]
 * (fclaw46_sync.c), which is the reference's per-face fm/fp/gm/gp and
 * edge_fluxes(., ., 0:1) arrays side by side.
 */
#include "fclaw46_oracle.h"

#define Q(a, i, j, m) (a)[((size_t) (m) * (my + 2 * mbc) + (size_t) ((j) + mbc - 1)) * (mx + 2 * mbc) + (i) + mbc - 1]

size_t oracle_ghost_pack_elems(int mx, int my, int mbc, int meqn, int refratio, int packarea,
                               int packregisters)
{
    const int mint = refratio * mbc;
    const long wg = (long) (2 * mbc + mx) * (2 * mbc + my);
    const long hole = (long) (mx - 2 * mint) * (my - 2 * mint);
    const long frsize = packregisters ? 2L * (4 * meqn + 2) * (mx + my) : 0;
    return (size_t) ((wg - hole) * (meqn + packarea) + frsize);
}

/* One sweep over the four strips of the band (the Fortran loop order), for
 * one field; copies q <-> pack and returns the number of values moved. */
static long band(int mx, int my, int mbc, int mint, double *q, int m, double *pack, int unpack)
{
    const int ng = mbc;
    long k = 0;
    int i, j;
#define MOVE(i_, j_) do { if (unpack) Q(q, i_, j_, m) = pack[k]; else pack[k] = Q(q, i_, j_, m); k++; } while (0)
    for (j = 1 - ng; j <= my - mint; j++)
        for (i = 1 - ng; i <= mint; i++) MOVE(i, j);
    for (j = 1 - ng; j <= mint; j++)
        for (i = mint + 1; i <= mx + ng; i++) MOVE(i, j);
    for (j = mint + 1; j <= my + ng; j++)
        for (i = mx - mint + 1; i <= mx + ng; i++) MOVE(i, j);
    for (j = my - mint + 1; j <= my + ng; j++)
        for (i = 1 - ng; i <= mx - mint; i++) MOVE(i, j);
#undef MOVE
    return k;
}

/* packmode 0 pack q, 1 unpack q, 2 pack q then area, 3 unpack both
 * (area: the reference's area(-mbc:mx+mbc+1, -mbc:my+mbc+1) array).
 * Returns the number of values moved, or -1 for a bad packmode. */
long oracle_local_ghost_pack(int mx, int my, int mbc, int meqn, int mint, double *q,
                             double *area, double *pack, int packmode)
{
    const int unpack = packmode == 1 || packmode == 3;
    long k = 0;
    int m;
    if (packmode < 0 || packmode > 3) return -1;
    for (m = 0; m < meqn; m++) k += band(mx, my, mbc, mint, q, m, pack + k, unpack);
    if (packmode >= 2) {
        /* the area array is one cell wider on every side: shift it into a
         * q-shaped view with mbc + 1 ghost layers */
        const int ma = mbc + 1;
        const int ng = mbc;
        int i, j;
#define A(i_, j_) area[(size_t) ((j_) + ma - 1) * (mx + 2 * ma) + (i_) + ma - 1]
#define MOVEA(i_, j_) do { if (unpack) A(i_, j_) = pack[k]; else pack[k] = A(i_, j_); k++; } while (0)
        for (j = 1 - ng; j <= my - mint; j++)
            for (i = 1 - ng; i <= mint; i++) MOVEA(i, j);
        for (j = 1 - ng; j <= mint; j++)
            for (i = mint + 1; i <= mx + ng; i++) MOVEA(i, j);
        for (j = mint + 1; j <= my + ng; j++)
            for (i = mx - mint + 1; i <= mx + ng; i++) MOVEA(i, j);
        for (j = my - mint + 1; j <= my + ng; j++)
            for (i = 1 - ng; i <= mx - mint; i++) MOVEA(i, j);
#undef MOVEA
#undef A
    }
    return k;
}

/* The registers of one patch into (unpack = 0) or out of (unpack = 1) the
 * message: for each face, meqn * len quadruples (fm, fp, edge flux 0, edge
 * flux 1), then len pairs (edge length, area).  Unpacking restores the
 * registers only (edge lengths and areas follow from dx, dy).  Returns the
 * number of values. */
long oracle_register_pack(int mx, int my, int meqn, double dx, double dy, double *reg,
                          double *pack, int unpack)
{
    long cnt = 0;
    int k, j;
    for (k = 0; k < 4; k++) {
        const int len = k < 2 ? my : mx;
        double *fm = reg + ((size_t) (k * 4 + ORACLE_REG_FM) * meqn) * mx;
        double *fp = reg + ((size_t) (k * 4 + ORACLE_REG_FP) * meqn) * mx;
        double *e0 = reg + ((size_t) (k * 4 + ORACLE_REG_EFIN) * meqn) * mx;
        double *e1 = reg + ((size_t) (k * 4 + ORACLE_REG_EFOUT) * meqn) * mx;
        for (j = 0; j < meqn * len; j++) {
            if (!unpack) {
                pack[cnt++] = fm[j];
                pack[cnt++] = fp[j];
                pack[cnt++] = e0[j];
                pack[cnt++] = e1[j];
            } else {
                fm[j] = pack[cnt++];
                fp[j] = pack[cnt++];
                e0[j] = pack[cnt++];
                e1[j] = pack[cnt++];
            }
        }
        for (j = 0; j < len; j++) {
            if (!unpack) {
                pack[cnt++] = k < 2 ? dy : dx;
                pack[cnt++] = dx * dy;
            } else {
                cnt += 2;
            }
        }
    }
    return cnt;
}
