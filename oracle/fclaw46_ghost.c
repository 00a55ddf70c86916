/*
 * Oracle (test infrastructure only): plain-C restatement of the reference's
 * clawpatch46 ghost-fill kernels, the default physical BC, the Kahan mass
 * diagnostic and the single-process ghost-fill orchestration.  Never linked
 * into the product.
 *
 *   copy        : src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_copy.f:17-133
 *   average     : src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_average.f:16-272
 *   interpolate : src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_interpolate.f:30-314
 *   slopes      : src/patches/clawpatch/fclaw2d_clawpatch_utils.f:11-66, 214-245
 *   transforms  : src/forestclaw2d.c:608-977, 1128-1380 (translations only: every
 *                 forest here is a brick of unit blocks with identical orientation)
 *   bc2         : src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_bc2_default.f:1-213
 *                 with mthbc = -1 on faces that are not physical (fc2d_clawpack46.cpp:121-128)
 *   order       : src/fclaw2d_ghost_fill.c:611-830 (one process: only the
 *                 INTERIOR_ONLY pass and the final physical-BC sweep act)
 *   face fill   : src/fclaw2d_face_neighbors.c:195-450
 *   corner fill : src/fclaw2d_corner_neighbors.c:40-560
 *   mass        : src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_diagnostics.f:13-55
 *   advance     : src/fclaw2d_advance.c:247-357 with subcycle = F (all levels step
 *                 with the same dt, then one ghost update over all levels)
 */
#include "fclaw46_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define Q(a, i, j, m) (a)[((size_t) (m) * (my + 2 * mbc) + (size_t) ((j) + mbc - 1)) * (mx + 2 * mbc) + (i) + mbc - 1]

/* ---------------- copy: fclaw2d_clawpatch46_copy.f ---------------- */

void oracle_copy_face(int mx, int my, int mbc, int meqn, double *qthis,
                      const double *qnbr, int iface, int ti, int tj)
{
    int mq, i1, j1;
    for (mq = 0; mq < meqn; mq++) {
        if (iface == 0) {
            for (j1 = 1; j1 <= my; j1++)
                for (i1 = 1 - mbc; i1 <= 0; i1++) Q(qthis, i1, j1, mq) = Q(qnbr, i1 + ti, j1 + tj, mq);
        } else if (iface == 1) {
            for (j1 = 1; j1 <= my; j1++)
                for (i1 = mx + 1; i1 <= mx + mbc; i1++) Q(qthis, i1, j1, mq) = Q(qnbr, i1 + ti, j1 + tj, mq);
        } else if (iface == 2) {
            for (j1 = 1 - mbc; j1 <= 0; j1++)
                for (i1 = 1; i1 <= mx; i1++) Q(qthis, i1, j1, mq) = Q(qnbr, i1 + ti, j1 + tj, mq);
        } else {
            for (j1 = my + 1; j1 <= my + mbc; j1++)
                for (i1 = 1; i1 <= mx; i1++) Q(qthis, i1, j1, mq) = Q(qnbr, i1 + ti, j1 + tj, mq);
        }
    }
}

static void corner_cell(int mx, int my, int icorner, int ibc, int jbc, int *i1, int *j1)
{
    if (icorner == 0) { *i1 = 1 - ibc; *j1 = 1 - jbc; }
    else if (icorner == 1) { *i1 = mx + ibc; *j1 = 1 - jbc; }
    else if (icorner == 2) { *i1 = 1 - ibc; *j1 = my + jbc; }
    else { *i1 = mx + ibc; *j1 = my + jbc; }
}

void oracle_copy_corner(int mx, int my, int mbc, int meqn, double *qthis,
                        const double *qnbr, int icorner, int ti, int tj)
{
    int mq, ibc, jbc, i1, j1;
    for (mq = 0; mq < meqn; mq++)
        for (ibc = 1; ibc <= mbc; ibc++)
            for (jbc = 1; jbc <= mbc; jbc++) {
                corner_cell(mx, my, icorner, ibc, jbc, &i1, &j1);
                Q(qthis, i1, j1, mq) = Q(qnbr, i1 + ti, j1 + tj, mq);
            }
}

/* transform_face_half / transform_corner_half for a translation:
 * children of coarse (ic,jc) in canonical order m = 2*ky + kx. */
static void half_children(int ic, int jc, int ti, int tj, int i2[4], int j2[4])
{
    const int di = 2 * ic - 1 + ti, dj = 2 * jc - 1 + tj;
    i2[0] = di;     j2[0] = dj;
    i2[1] = di + 1; j2[1] = dj;
    i2[2] = di;     j2[2] = dj + 1;
    i2[3] = di + 1; j2[3] = dj + 1;
}

/* ---------------- average: fclaw2d_clawpatch46_average.f (manifold = 0) ---------------- */

void oracle_average_face(int mx, int my, int mbc, int meqn, double *qcoarse,
                         const double *qfine, int idir, int iface_coarse,
                         int ti, int tj)
{
    int mq, ic, jc, ibc, jbc, m, i2[4], j2[4];
    for (mq = 0; mq < meqn; mq++) {
        if (idir == 0) {
            for (jc = 1; jc <= my; jc++)
                for (ibc = 1; ibc <= mbc; ibc++) {
                    int skip = 0;
                    double sum;
                    ic = (iface_coarse == 0) ? 1 - ibc : mx + ibc;
                    half_children(ic, jc, ti, tj, i2, j2);
                    for (m = 0; m < 4; m++)
                        if (!(1 <= i2[m] && i2[m] <= mx && 1 <= j2[m] && j2[m] <= my)) { skip = 1; break; }
                    if (skip) continue;
                    sum = 0;
                    for (m = 0; m < 4; m++) sum = sum + Q(qfine, i2[m], j2[m], mq);
                    Q(qcoarse, ic, jc, mq) = sum / 4.0;
                }
        } else {
            for (jbc = 1; jbc <= mbc; jbc++)
                for (ic = 1; ic <= mx; ic++) {
                    int skip = 0;
                    double sum;
                    jc = (iface_coarse == 2) ? 1 - jbc : my + jbc;
                    half_children(ic, jc, ti, tj, i2, j2);
                    for (m = 0; m < 4; m++)
                        if (!(1 <= i2[m] && i2[m] <= mx && 1 <= j2[m] && j2[m] <= my)) skip = 1;
                    if (skip) continue;
                    sum = 0;
                    for (m = 0; m < 4; m++) sum = sum + Q(qfine, i2[m], j2[m], mq);
                    Q(qcoarse, ic, jc, mq) = sum / 4.0;
                }
        }
    }
}

void oracle_average_corner(int mx, int my, int mbc, int meqn, double *qcoarse,
                           const double *qfine, int icorner, int ti, int tj)
{
    int ibc, jbc, mq, m, i1, j1, i2[4], j2[4];
    for (ibc = 1; ibc <= mbc; ibc++)
        for (jbc = 1; jbc <= mbc; jbc++) {
            corner_cell(mx, my, icorner, ibc, jbc, &i1, &j1);
            half_children(i1, j1, ti, tj, i2, j2);
            for (mq = 0; mq < meqn; mq++) {
                double sum = 0;
                for (m = 0; m < 4; m++) sum = sum + Q(qfine, i2[m], j2[m], mq);
                Q(qcoarse, i1, j1, mq) = sum / 4.0;
            }
        }
}

/* ---------------- interpolate: fclaw2d_clawpatch46_interpolate.f ---------------- */

/* fclaw2d_clawpatch_compute_slopes with mth = 5 (AMRClaw slopes). */
double oracle_compute_slopes(double sl, double sr)
{
    double sc = (sl + sr) / 2.0;
    double mn = fmin(fmin(2 * fabs(sl), 2 * fabs(sr)), fabs(sc));
    double sgn_prod = copysign(1.0, sl * sr);
    return mn * fmax(0.0, sgn_prod) * copysign(1.0, sc);
}

static void interp_cell(int mx, int my, int mbc, const double *qcoarse, double *qfine,
                        int mq, int ic, int jc, const int i2[4], const int j2[4])
{
    /* build_transform gives a = 2*I for a translation, so df(:,m) = (ii, jj) */
    static const int dfi[4] = {0, 1, 0, 1}, dfj[4] = {0, 0, 1, 1};
    static const double shx[4] = {-0.25, 0.25, -0.25, 0.25}, shy[4] = {-0.25, -0.25, 0.25, 0.25};
    double qc = Q(qcoarse, ic, jc, mq), sl, sr, gradx, grady;
    int m;
    sl = (qc - Q(qcoarse, ic - 1, jc, mq));
    sr = (Q(qcoarse, ic + 1, jc, mq) - qc);
    gradx = oracle_compute_slopes(sl, sr);
    sl = (qc - Q(qcoarse, ic, jc - 1, mq));
    sr = (Q(qcoarse, ic, jc + 1, mq) - qc);
    grady = oracle_compute_slopes(sl, sr);
    for (m = 0; m < 4; m++) {
        int iff = i2[0] + dfi[m], jff = j2[0] + dfj[m];
        Q(qfine, iff, jff, mq) = qc + gradx * shx[m] + grady * shy[m];
    }
}

void oracle_interpolate_face(int mx, int my, int mbc, int meqn,
                             const double *qcoarse, double *qfine, int idir,
                             int iface_coarse, int ti, int tj)
{
    int mq, ibc, jbc, ic, jc, m, i2[4], j2[4];
    for (mq = 0; mq < meqn; mq++) {
        if (idir == 0) {
            for (ibc = 1; ibc <= mbc / 2; ibc++) {
                ic = (iface_coarse == 0) ? ibc : mx - ibc + 1;
                for (jc = 1; jc <= mx; jc++) { /* the reference bounds jc by mx */
                    int skip = 0;
                    half_children(ic, jc, ti, tj, i2, j2);
                    for (m = 0; m < 4; m++)
                        if (!(1 - mbc <= i2[m] && i2[m] <= mx + mbc && 1 - mbc <= j2[m] && j2[m] <= my + mbc)) { skip = 1; break; }
                    if (!skip) interp_cell(mx, my, mbc, qcoarse, qfine, mq, ic, jc, i2, j2);
                }
            }
        } else {
            for (jbc = 1; jbc <= mbc / 2; jbc++) {
                jc = (iface_coarse == 2) ? jbc : my - jbc + 1;
                for (ic = 1; ic <= mx; ic++) {
                    int skip = 0;
                    half_children(ic, jc, ti, tj, i2, j2);
                    for (m = 0; m < 4; m++)
                        if (!(1 - mbc <= i2[m] && i2[m] <= mx + mbc && 1 - mbc <= j2[m] && j2[m] <= my + mbc)) { skip = 1; break; }
                    if (!skip) interp_cell(mx, my, mbc, qcoarse, qfine, mq, ic, jc, i2, j2);
                }
            }
        }
    }
}

void oracle_interpolate_corner(int mx, int my, int mbc, int meqn,
                               const double *qcoarse, double *qfine,
                               int icorner, int ti, int tj)
{
    int ibc, jbc, ic, jc, mq, i2[4], j2[4];
    for (ibc = 1; ibc <= mbc / 2; ibc++)
        for (jbc = 1; jbc <= mbc / 2; jbc++) {
            if (icorner == 0) { ic = ibc; jc = jbc; }
            else if (icorner == 1) { ic = mx - ibc + 1; jc = jbc; }
            else if (icorner == 2) { ic = ibc; jc = my - jbc + 1; }
            else { ic = mx - ibc + 1; jc = my - jbc + 1; }
            half_children(ic, jc, ti, tj, i2, j2);
            for (mq = 0; mq < meqn; mq++) interp_cell(mx, my, mbc, qcoarse, qfine, mq, ic, jc, i2, j2);
        }
}

/* ---------------- bc2: clawpack46_bc2_default.f ---------------- */

void oracle_bc2(int mx, int my, int mbc, int meqn, double *q, const int *mthbc)
{
    int m, i, j, ibc, jbc;
    /* left */
    if (mthbc[0] == 1) {
        for (m = 0; m < meqn; m++) for (ibc = 1; ibc <= mbc; ibc++) for (j = 1 - mbc; j <= my + mbc; j++)
            Q(q, 1 - ibc, j, m) = Q(q, 1, j, m);
    } else if (mthbc[0] == 3) {
        for (m = 0; m < meqn; m++) for (ibc = 1; ibc <= mbc; ibc++) for (j = 1 - mbc; j <= my + mbc; j++)
            Q(q, 1 - ibc, j, m) = Q(q, ibc, j, m);
        for (ibc = 1; ibc <= mbc; ibc++) for (j = 1 - mbc; j <= my + mbc; j++)
            Q(q, 1 - ibc, j, 1) = -Q(q, ibc, j, 1);
    }
    /* right */
    if (mthbc[1] == 1) {
        for (m = 0; m < meqn; m++) for (ibc = 1; ibc <= mbc; ibc++) for (j = 1 - mbc; j <= my + mbc; j++)
            Q(q, mx + ibc, j, m) = Q(q, mx, j, m);
    } else if (mthbc[1] == 3) {
        for (m = 0; m < meqn; m++) for (ibc = 1; ibc <= mbc; ibc++) for (j = 1 - mbc; j <= my + mbc; j++)
            Q(q, mx + ibc, j, m) = Q(q, mx + 1 - ibc, j, m);
        for (ibc = 1; ibc <= mbc; ibc++) for (j = 1 - mbc; j <= my + mbc; j++)
            Q(q, mx + ibc, j, 1) = -Q(q, mx + 1 - ibc, j, 1);
    }
    /* bottom */
    if (mthbc[2] == 1) {
        for (m = 0; m < meqn; m++) for (jbc = 1; jbc <= mbc; jbc++) for (i = 1 - mbc; i <= mx + mbc; i++)
            Q(q, i, 1 - jbc, m) = Q(q, i, 1, m);
    } else if (mthbc[2] == 3) {
        for (m = 0; m < meqn; m++) for (jbc = 1; jbc <= mbc; jbc++) for (i = 1 - mbc; i <= mx + mbc; i++)
            Q(q, i, 1 - jbc, m) = Q(q, i, jbc, m);
        for (jbc = 1; jbc <= mbc; jbc++) for (i = 1 - mbc; i <= mx + mbc; i++)
            Q(q, i, 1 - jbc, 2) = -Q(q, i, jbc, 2);
    }
    /* top */
    if (mthbc[3] == 1) {
        for (m = 0; m < meqn; m++) for (jbc = 1; jbc <= mbc; jbc++) for (i = 1 - mbc; i <= mx + mbc; i++)
            Q(q, i, my + jbc, m) = Q(q, i, my, m);
    } else if (mthbc[3] == 3) {
        for (m = 0; m < meqn; m++) for (jbc = 1; jbc <= mbc; jbc++) for (i = 1 - mbc; i <= mx + mbc; i++)
            Q(q, i, my + jbc, m) = Q(q, i, my + 1 - jbc, m);
        for (jbc = 1; jbc <= mbc; jbc++) for (i = 1 - mbc; i <= mx + mbc; i++)
            Q(q, i, my + jbc, 2) = -Q(q, i, my + 1 - jbc, 2);
    }
}

/* ---------------- forest neighbours (p4est semantics, brick of unit blocks) ---------------- */

enum { NB_NONE = 0, NB_SAME = 1, NB_HALF = 2, NB_DOUBLE = 3, NB_BOUNDARY = 4 };

typedef struct ohash {
    size_t cap;
    long long *key;
    int *val;
} ohash_t;

static long long hkey(int level, long long gx, long long gy)
{
    return ((long long) level << 56) ^ (gx << 28) ^ gy;
}

static void hash_init(ohash_t *h, int n)
{
    size_t cap = 16;
    while (cap < (size_t) n * 4) cap <<= 1;
    h->cap = cap;
    h->key = (long long *) malloc(cap * sizeof(long long));
    h->val = (int *) malloc(cap * sizeof(int));
    memset(h->val, -1, cap * sizeof(int));
}

static void hash_put(ohash_t *h, long long k, int v)
{
    size_t p = ((unsigned long long) k * 0x9E3779B97F4A7C15ull) & (h->cap - 1);
    while (h->val[p] >= 0) p = (p + 1) & (h->cap - 1);
    h->key[p] = k;
    h->val[p] = v;
}

static int hash_get(const ohash_t *h, long long k)
{
    size_t p = ((unsigned long long) k * 0x9E3779B97F4A7C15ull) & (h->cap - 1);
    while (h->val[p] >= 0) {
        if (h->key[p] == k) return h->val[p];
        p = (p + 1) & (h->cap - 1);
    }
    return -1;
}

static void hash_free(ohash_t *h) { free(h->key); free(h->val); }

/* Wrap a patch coordinate at `level`; returns 0 if outside a non-periodic domain. */
static int wrap_coord(const oracle_domain_t *d, int level, long long *gx, long long *gy)
{
    const long long nx = (long long) d->nbx << level, ny = (long long) d->nby << level;
    if (*gx < 0 || *gx >= nx) {
        if (!d->periodic_x) return 0;
        *gx = ((*gx % nx) + nx) % nx;
    }
    if (*gy < 0 || *gy >= ny) {
        if (!d->periodic_y) return 0;
        *gy = ((*gy % ny) + ny) % ny;
    }
    return 1;
}

static int find_leaf(const oracle_domain_t *d, const ohash_t *h, int level, long long gx, long long gy)
{
    if (level < 0) return -1;
    if (!wrap_coord(d, level, &gx, &gy)) return -1;
    return hash_get(h, hkey(level, gx, gy));
}

/* Physical-boundary flags of patch p, faces 0..3 (fclaw2d_physical_get_bc). */
static void phys_bdry(const oracle_domain_t *d, int p, int bdry[4])
{
    const int L = d->level[p];
    const long long nx = (long long) d->nbx << L, ny = (long long) d->nby << L;
    bdry[0] = !d->periodic_x && d->gx[p] == 0;
    bdry[1] = !d->periodic_x && d->gx[p] == nx - 1;
    bdry[2] = !d->periodic_y && d->gy[p] == 0;
    bdry[3] = !d->periodic_y && d->gy[p] == ny - 1;
}

typedef struct face_nb {
    int type;
    int nb[2];         /* neighbour patch(es); two for NB_HALF */
    int ti[2], tj[2];  /* index translation (see oracle_* kernels) */
} face_nb_t;

static const int face_dx[4] = {-1, 1, 0, 0}, face_dy[4] = {0, 0, -1, 1};
static const int corner_dx[4] = {-1, 1, -1, 1}, corner_dy[4] = {-1, -1, 1, 1};

/* fclaw2d_patch_face_neighbors for this brick, with translations. */
static void face_neighbor(const oracle_domain_t *d, const ohash_t *h, int p, int iface, face_nb_t *f)
{
    const int L = d->level[p], mx = d->mx, my = d->my;
    const long long gx = d->gx[p], gy = d->gy[p];
    const long long nx = gx + face_dx[iface], ny = gy + face_dy[iface];
    int bdry[4], q, k;
    phys_bdry(d, p, bdry);
    f->type = NB_NONE;
    if (bdry[iface]) { f->type = NB_BOUNDARY; return; }
    /* same size: neighbour index = this index - (offset)*m */
    {
        long long wx = nx, wy = ny;
        if (!wrap_coord(d, L, &wx, &wy)) { f->type = NB_BOUNDARY; return; }
        q = hash_get(h, hkey(L, wx, wy));
        if (q >= 0) {
            f->type = NB_SAME; f->nb[0] = q;
            f->ti[0] = -face_dx[iface] * mx; f->tj[0] = -face_dy[iface] * my;
            return;
        }
        q = (L >= 1) ? hash_get(h, hkey(L - 1, wx >> 1, wy >> 1)) : -1;
        if (q >= 0) { f->type = NB_DOUBLE; f->nb[0] = q; return; }
    }
    /* half size: two children adjacent to this face */
    f->type = NB_HALF;
    for (k = 0; k < 2; k++) {
        long long cx, cy;
        if (iface == 0) { cx = 2 * nx + 1; cy = 2 * gy + k; }
        else if (iface == 1) { cx = 2 * nx; cy = 2 * gy + k; }
        else if (iface == 2) { cx = 2 * gx + k; cy = 2 * ny + 1; }
        else { cx = 2 * gx + k; cy = 2 * ny; }
        q = find_leaf(d, h, L + 1, cx, cy);
        f->nb[k] = q;
        /* first fine child of coarse (ic,jc): 2*ic-1 + (2*gx - cx)*mx (unwrapped cx) */
        f->ti[k] = (int) ((2 * gx - cx) * mx);
        f->tj[k] = (int) ((2 * gy - cy) * my);
    }
    if (f->nb[0] < 0 || f->nb[1] < 0) f->type = NB_NONE; /* unbalanced forest */
}

/* fclaw2d_patch_corner_neighbors for this brick (get_corner_neighbor cases 1-3). */
static void corner_neighbor(const oracle_domain_t *d, const ohash_t *h, int p, int icorner, face_nb_t *f)
{
    const int L = d->level[p], mx = d->mx, my = d->my;
    const long long gx = d->gx[p], gy = d->gy[p];
    const long long nx = gx + corner_dx[icorner], ny = gy + corner_dy[icorner];
    /* corner point in level-L lattice coordinates */
    const long long px = gx + (corner_dx[icorner] > 0), py = gy + (corner_dy[icorner] > 0);
    int q;
    f->type = NB_NONE;
    q = find_leaf(d, h, L, nx, ny);
    if (q >= 0) {
        f->type = NB_SAME; f->nb[0] = q;
        f->ti[0] = -corner_dx[icorner] * mx; f->tj[0] = -corner_dy[icorner] * my;
        return;
    }
    if (L >= 1) {
        /* coarser leaf covering the diagonal cell: a corner neighbour only if the
         * corner point is also one of its corners (else this is a hanging node) */
        long long wx = nx, wy = ny, cx, cy;
        if (wrap_coord(d, L, &wx, &wy)) {
            cx = wx >> 1; cy = wy >> 1;
            q = find_leaf(d, h, L - 1, cx, cy);
            if (q >= 0) {
                /* corner point relative to the unwrapped coarse cell */
                long long ux = (nx >> 1), uy = (ny >> 1);
                long long rx = px - 2 * ux, ry = py - 2 * uy;
                if ((rx == 0 || rx == 2) && (ry == 0 || ry == 2)) {
                    f->type = NB_DOUBLE; f->nb[0] = q;
                }
                return;
            }
        }
    }
    {
        long long cx = 2 * nx + (corner_dx[icorner] < 0 ? 1 : 0);
        long long cy = 2 * ny + (corner_dy[icorner] < 0 ? 1 : 0);
        q = find_leaf(d, h, L + 1, cx, cy);
        if (q >= 0) {
            f->type = NB_HALF; f->nb[0] = q;
            f->ti[0] = (int) ((2 * gx - cx) * mx);
            f->tj[0] = (int) ((2 * gy - cy) * my);
        }
    }
}

/* get_corner_type: interior corner iff neither adjoining face is physical. */
static int corner_is_interior(const oracle_domain_t *d, int p, int icorner)
{
    static const int cf[4][2] = {{0, 2}, {1, 2}, {0, 3}, {1, 3}};
    int bdry[4];
    phys_bdry(d, p, bdry);
    return !bdry[cf[icorner][0]] && !bdry[cf[icorner][1]];
}

#define PATCH(qall, p) ((qall) + (size_t) (p) * psize)

enum { FILL_AVERAGE, FILL_COPY, FILL_INTERP };

/* One level of face exchanges: `qt` holds the patches of `level` (coarse side
 * for average/interpolate), `qn` their neighbours (fine side, or same level).
 * With time interpolation the coarse level's arrays are its time-interpolated
 * copies (fclaw2d_clawpatch_timesync_data, src/patches/clawpatch/fclaw2d_clawpatch.cpp:1746-1754). */
static void face_fill_level(const oracle_domain_t *d, const ohash_t *h, double *qt, double *qn,
                            int level, int filltype)
{
    const int mx = d->mx, my = d->my, mbc = d->mbc, meqn = d->meqn;
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    int p, iface, ig;
    for (p = 0; p < d->npatch; p++) {
        if (d->level[p] != level) continue;
        for (iface = 0; iface < 4; iface++) {
            face_nb_t f;
            face_neighbor(d, h, p, iface, &f);
            if (f.type == NB_HALF) {
                for (ig = 0; ig < 2; ig++) {
                    if (filltype == FILL_INTERP)
                        oracle_interpolate_face(mx, my, mbc, meqn, PATCH(qt, p), PATCH(qn, f.nb[ig]),
                                                iface / 2, iface, f.ti[ig], f.tj[ig]);
                    else if (filltype == FILL_AVERAGE)
                        oracle_average_face(mx, my, mbc, meqn, PATCH(qt, p), PATCH(qn, f.nb[ig]),
                                            iface / 2, iface, f.ti[ig], f.tj[ig]);
                }
            } else if (f.type == NB_SAME && filltype == FILL_COPY) {
                oracle_copy_face(mx, my, mbc, meqn, PATCH(qt, p), PATCH(qn, f.nb[0]), iface, f.ti[0], f.tj[0]);
            }
        }
    }
}

static void corner_fill_level(const oracle_domain_t *d, const ohash_t *h, double *qt,
                              double *qn, int level, int filltype)
{
    const int mx = d->mx, my = d->my, mbc = d->mbc, meqn = d->meqn;
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    int p, ic;
    for (p = 0; p < d->npatch; p++) {
        if (d->level[p] != level) continue;
        for (ic = 0; ic < 4; ic++) {
            face_nb_t f;
            if (!corner_is_interior(d, p, ic)) continue;
            corner_neighbor(d, h, p, ic, &f);
            if (f.type == NB_HALF) {
                if (filltype == FILL_INTERP)
                    oracle_interpolate_corner(mx, my, mbc, meqn, PATCH(qt, p), PATCH(qn, f.nb[0]), ic, f.ti[0], f.tj[0]);
                else if (filltype == FILL_AVERAGE)
                    oracle_average_corner(mx, my, mbc, meqn, PATCH(qt, p), PATCH(qn, f.nb[0]), ic, f.ti[0], f.tj[0]);
            } else if (f.type == NB_SAME && filltype == FILL_COPY) {
                oracle_copy_corner(mx, my, mbc, meqn, PATCH(qt, p), PATCH(qn, f.nb[0]), ic, f.ti[0], f.tj[0]);
            }
        }
    }
}

static void physbc_level(const oracle_domain_t *d, double *q, int level)
{
    const int mx = d->mx, my = d->my, mbc = d->mbc, meqn = d->meqn;
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    int p, k;
    for (p = 0; p < d->npatch; p++) {
        int bdry[4], mthbc[4];
        if (d->level[p] != level) continue;
        phys_bdry(d, p, bdry);
        if (!(bdry[0] || bdry[1] || bdry[2] || bdry[3])) continue;
        for (k = 0; k < 4; k++) mthbc[k] = bdry[k] ? d->mthbc[k] : -1;
        oracle_bc2(mx, my, mbc, meqn, PATCH(q, p), mthbc);
    }
}

/* fclaw2d_ghost_update_async over levels minlevel..maxlevel with one process
 * (the INTERIOR_ONLY pass, then the final physical-BC sweep).  With
 * time_interp, level minlevel-1 takes part through its time-interpolated
 * arrays qti: fine level minlevel averages into their ghost cells, they copy
 * among themselves, get physical BCs, and interpolate to level minlevel's
 * ghost cells (src/fclaw2d_ghost_fill.c:341-442).  The reference also applies
 * physical BCs to time-interpolated arrays that nothing reads afterwards
 * (level minlevel in the fine sweep, level minlevel-1 in the final sweep);
 * those are left out, since q does not depend on them. */
static void ghost_update_range(const oracle_domain_t *d, const ohash_t *h, double *q, double *qti,
                               int minlevel, int maxlevel, int time_interp)
{
    const int ti = minlevel - 1, mincoarse = minlevel, maxcoarse = maxlevel - 1;
    int level;
    /* average_fine2coarse_ghost */
    for (level = maxcoarse; level >= mincoarse; level--) {
        face_fill_level(d, h, q, q, level, FILL_AVERAGE);
        corner_fill_level(d, h, q, q, level, FILL_AVERAGE);
    }
    if (time_interp) {
        face_fill_level(d, h, qti, q, ti, FILL_AVERAGE);
        corner_fill_level(d, h, qti, q, ti, FILL_AVERAGE);
    }
    /* copy_samelevel */
    for (level = maxlevel; level >= minlevel; level--) {
        face_fill_level(d, h, q, q, level, FILL_COPY);
        corner_fill_level(d, h, q, q, level, FILL_COPY);
    }
    if (time_interp) {
        face_fill_level(d, h, qti, qti, ti, FILL_COPY);
        corner_fill_level(d, h, qti, qti, ti, FILL_COPY);
    }
    /* fill_physical_ghost on coarse levels */
    for (level = maxcoarse; level >= mincoarse; level--) physbc_level(d, q, level);
    if (time_interp) physbc_level(d, qti, ti);
    /* interpolate_coarse2fine_ghost */
    for (level = maxcoarse; level >= mincoarse; level--) {
        face_fill_level(d, h, q, q, level, FILL_INTERP);
        corner_fill_level(d, h, q, q, level, FILL_INTERP);
    }
    if (time_interp) {
        face_fill_level(d, h, qti, q, ti, FILL_INTERP);
        corner_fill_level(d, h, qti, q, ti, FILL_INTERP);
    }
    /* fill_physical_ghost on fine levels, then on all levels */
    for (level = maxlevel; level >= mincoarse + 1; level--) physbc_level(d, q, level);
    for (level = maxlevel; level >= minlevel; level--) physbc_level(d, q, level);
}

static void domain_levels(const oracle_domain_t *d, ohash_t *h, int *minlevel, int *maxlevel)
{
    int p;
    *minlevel = 1 << 30;
    *maxlevel = -1;
    hash_init(h, d->npatch);
    for (p = 0; p < d->npatch; p++) {
        hash_put(h, hkey(d->level[p], d->gx[p], d->gy[p]), p);
        if (d->level[p] < *minlevel) *minlevel = d->level[p];
        if (d->level[p] > *maxlevel) *maxlevel = d->level[p];
    }
}

int oracle_ghost_update(const oracle_domain_t *d, double *q)
{
    ohash_t h;
    int minlevel, maxlevel;
    if (d->npatch <= 0) return 0;
    domain_levels(d, &h, &minlevel, &maxlevel);
    ghost_update_range(d, &h, q, NULL, minlevel, maxlevel, 0);
    hash_free(&h);
    return 0;
}

double oracle_update_all(const oracle_domain_t *d, const oracle_claw_t *claw,
                         double *q, const double *aux, double dt)
{
    const int mx = d->mx, my = d->my, mbc = d->mbc, meqn = d->meqn;
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    const size_t asize = (size_t) (d->maux > 0 ? d->maux : 0) * (mx + 2 * mbc) * (my + 2 * mbc);
    const double wx = (d->bx - d->ax) / d->nbx, wy = (d->by - d->ay) / d->nby;
    double maxcfl = 0.0;
    int p;
    for (p = 0; p < d->npatch; p++) {
        const double nl = (double) (1 << d->level[p]);
        const double dx = wx / (nl * mx), dy = wy / (nl * my);
        double cfl;
        oracle_step2_patch(claw, mx, my, mbc, PATCH(q, p), aux ? aux + (size_t) p * asize : NULL,
                           dx, dy, dt, &cfl, NULL, NULL, NULL, NULL);
        maxcfl = fmax(cfl, maxcfl);
    }
    return maxcfl;
}

double oracle_advance(const oracle_domain_t *d, const oracle_claw_t *claw,
                      double *q, const double *aux, double dt, int nsteps)
{
    return oracle_advance_ts(d, claw, q, aux, dt, nsteps, 0);
}

/* Zero the registers of one face: fclaw2d_clawpatch_time_sync_reset
 * (src/patches/clawpatch/fclaw2d_clawpatch_conservation.c:207-278). */
static void reset_face(double *reg, int mx, int meqn, int face)
{
    int kind, e;
    for (kind = 0; kind < 4; kind++)
        for (e = 0; e < meqn * mx; e++) reg[((size_t) face * 4 + kind) * meqn * mx + e] = 0.0;
}

/* correct_coarse_cells (src/fclaw2d_time_sync.c:118-158) with one process, for
 * levels maxlevel..minlevel: per level, finest first, the same-size
 * corrections of every patch face (cb_face_fill with
 * FCLAW2D_TIME_SYNC_SAMESIZE, src/fclaw2d_face_neighbors.c:368-383) and the
 * reset of those registers, then, below the finest level, the fine-to-coarse
 * corrections of every coarse face with half-size neighbours (:321-336) and
 * the reset of the registers on both sides of those faces; finally the reset
 * of the registers on physical boundaries.  Registers of faces between level
 * minlevel and minlevel-1 keep accumulating (subcycling). */
static void time_sync_range(const oracle_domain_t *d, const ohash_t *h, double *q, double *reg,
                            int minlevel, int maxlevel)
{
    const int mx = d->mx, my = d->my, mbc = d->mbc, meqn = d->meqn;
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    const size_t rsize = oracle_sync_register_count(mx, meqn);
    const double wx = (d->bx - d->ax) / d->nbx, wy = (d->by - d->ay) / d->nby;
    int p, level, k, g;
    for (level = maxlevel; level >= minlevel; level--) {
        for (p = 0; p < d->npatch; p++) {
            const double nl = (double) (1 << d->level[p]);
            const double dx = wx / (nl * mx), dy = wy / (nl * my);
            if (d->level[p] != level) continue;
            for (k = 0; k < 4; k++) {
                face_nb_t f;
                face_neighbor(d, h, p, k, &f);
                if (f.type == NB_SAME)
                    oracle_time_sync_samesize(mx, my, mbc, meqn, PATCH(q, p), reg + p * rsize,
                                              reg + (size_t) f.nb[0] * rsize, k, dx, dy);
            }
        }
        for (p = 0; p < d->npatch; p++) {
            if (d->level[p] != level) continue;
            for (k = 0; k < 4; k++) {
                face_nb_t f;
                face_neighbor(d, h, p, k, &f);
                if (f.type == NB_SAME) reset_face(reg + p * rsize, mx, meqn, k);
            }
        }
        if (level == maxlevel) continue;
        for (p = 0; p < d->npatch; p++) {
            const double nl = (double) (1 << d->level[p]);
            const double dx = wx / (nl * mx), dy = wy / (nl * my);
            if (d->level[p] != level) continue;
            for (k = 0; k < 4; k++) {
                face_nb_t f;
                face_neighbor(d, h, p, k, &f);
                if (f.type != NB_HALF) continue;
                for (g = 0; g < 2; g++)
                    oracle_time_sync_f2c(mx, my, mbc, meqn, PATCH(q, p), reg + p * rsize,
                                         reg + (size_t) f.nb[g] * rsize, k, f.ti[g], f.tj[g], dx, dy);
            }
        }
        for (p = 0; p < d->npatch; p++) {
            if (d->level[p] != level && d->level[p] != level + 1) continue;
            for (k = 0; k < 4; k++) {
                face_nb_t f;
                face_neighbor(d, h, p, k, &f);
                if ((d->level[p] == level && f.type == NB_HALF) ||
                    (d->level[p] == level + 1 && f.type == NB_DOUBLE))
                    reset_face(reg + p * rsize, mx, meqn, k);
            }
        }
    }
    for (p = 0; p < d->npatch; p++) {
        if (d->level[p] < minlevel) continue;
        for (k = 0; k < 4; k++) {
            face_nb_t f;
            face_neighbor(d, h, p, k, &f);
            if (f.type == NB_BOUNDARY) reset_face(reg + p * rsize, mx, meqn, k);
        }
    }
}

int oracle_time_sync(const oracle_domain_t *d, double *q, double *reg)
{
    ohash_t h;
    int minlevel, maxlevel;
    if (d->npatch <= 0) return 0;
    if (d->mx != d->my) return -1;
    domain_levels(d, &h, &minlevel, &maxlevel);
    time_sync_range(d, &h, q, reg, minlevel, maxlevel);
    hash_free(&h);
    return 0;
}

/* One patch update with, if reg is given, the flux registers filled before
 * and after (fc2d_clawpack46.cpp:317-392); with qlast, the state is saved
 * first (fclaw2d_clawpatch_save_current_step before every update). */
static double update_patch(const oracle_domain_t *d, const oracle_claw_t *claw, double *q,
                           const double *aux, double *qlast, double *reg, double *fl, int p,
                           double dt)
{
    const int mx = d->mx, my = d->my, mbc = d->mbc, meqn = d->meqn;
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    const size_t asize = (size_t) (d->maux > 0 ? d->maux : 0) * (mx + 2 * mbc) * (my + 2 * mbc);
    const size_t rsize = oracle_sync_register_count(mx, meqn);
    const double wx = (d->bx - d->ax) / d->nbx, wy = (d->by - d->ay) / d->nby;
    const double nl = (double) (1 << d->level[p]);
    const double dx = wx / (nl * mx), dy = wy / (nl * my);
    const double *ap = aux ? aux + (size_t) p * asize : NULL;
    double cfl;
    if (qlast) memcpy(PATCH(qlast, p), PATCH(q, p), psize * sizeof(double));
    if (!reg) {
        oracle_step2_patch(claw, mx, my, mbc, PATCH(q, p), ap, dx, dy, dt, &cfl, NULL, NULL, NULL, NULL);
        return cfl;
    }
    oracle_store_flux(claw, mx, my, mbc, PATCH(q, p), ap, dt, dx, dy, reg + p * rsize);
    oracle_step2_patch(claw, mx, my, mbc, PATCH(q, p), ap, dx, dy, dt, &cfl,
                       fl, fl + psize, fl + 2 * psize, fl + 3 * psize);
    oracle_accumulate_waves(mx, my, mbc, meqn, dt, dx, dy, fl + psize, fl, fl + 3 * psize,
                            fl + 2 * psize, reg + p * rsize);
    return cfl;
}

/* nsteps of fclaw2d_advance_all_levels with subcycle = F
 * (src/fclaw2d_advance.c:247-357): every patch steps at dt (with, if
 * time_sync, the flux registers of fc2d_clawpack46.cpp:317-392 filled before
 * and after its step), then one ghost update over all levels, then, if
 * time_sync, fclaw2d_time_sync. */
double oracle_advance_ts(const oracle_domain_t *d, const oracle_claw_t *claw, double *q,
                         const double *aux, double dt, int nsteps, int time_sync)
{
    const int mx = d->mx, my = d->my, mbc = d->mbc, meqn = d->meqn;
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    const size_t rsize = oracle_sync_register_count(mx, meqn);
    double *fl = NULL, *reg = NULL, maxcfl = 0.0;
    int n, p;
    if (time_sync) {
        fl = (double *) malloc(4 * psize * sizeof(double));
        reg = (double *) calloc((size_t) d->npatch * rsize, sizeof(double));
    }
    for (n = 0; n < nsteps; n++) {
        for (p = 0; p < d->npatch; p++)
            maxcfl = fmax(update_patch(d, claw, q, aux, NULL, reg, fl, p, dt), maxcfl);
        oracle_ghost_update(d, q);
        if (time_sync) oracle_time_sync(d, q, reg);
    }
    free(fl);
    free(reg);
    return maxcfl;
}

/* ---------------- subcycling: fclaw2d_advance_all_levels with subcycle = T ----------------
 * src/fclaw2d_advance.c:54-357.  Level L steps 2^(L-minlevel) times with
 * dt/2^(L-minlevel); a coarser level steps first (advance_level recursion)
 * and its time-interpolated border band (fclaw2d_clawpatch46_fort_timeinterp,
 * src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_timeinterp.f:1-58, called
 * through clawpatch_setup_timeinterp, fclaw2d_clawpatch.cpp:474-526) feeds the
 * ghost cells of the finer levels at intermediate times. */
typedef struct sub_ctx {
    const oracle_domain_t *d;
    const oracle_claw_t *claw;
    ohash_t h;
    double *q, *qlast, *qti, *reg, *fl;
    const double *aux;
    int minlevel, maxlevel;
    int last_step[32], step_inc[32], total_steps[32];
    double dt_step[32], current_time[32];
} sub_ctx_t;

/* One patch: qinterp = qlast + alpha (qcurr - qlast) on the interior band of
 * width mint = mbc, in the four loops of fclaw2d_clawpatch46_fort_timeinterp
 * (fclaw2d_clawpatch46_timeinterp.f:28-62). */
void oracle_timeinterp_patch(int mx, int my, int mbc, int meqn, const double *qc,
                             const double *ql, double *qi, double alpha)
{
    const int mint = mbc;
    int i, j, m;
    for (m = 0; m < meqn; m++) {
        for (j = 1; j <= my - mint; j++)
            for (i = 1; i <= mint; i++)
                Q(qi, i, j, m) = Q(ql, i, j, m) + alpha * (Q(qc, i, j, m) - Q(ql, i, j, m));
        for (j = 1; j <= mint; j++)
            for (i = mint + 1; i <= mx; i++)
                Q(qi, i, j, m) = Q(ql, i, j, m) + alpha * (Q(qc, i, j, m) - Q(ql, i, j, m));
        for (j = mint + 1; j <= my; j++)
            for (i = mx - mint + 1; i <= mx; i++)
                Q(qi, i, j, m) = Q(ql, i, j, m) + alpha * (Q(qc, i, j, m) - Q(ql, i, j, m));
        for (j = my - mint + 1; j <= my; j++)
            for (i = 1; i <= mx - mint; i++)
                Q(qi, i, j, m) = Q(ql, i, j, m) + alpha * (Q(qc, i, j, m) - Q(ql, i, j, m));
    }
}

static void timeinterp_level(sub_ctx_t *c, int level, double alpha)
{
    const oracle_domain_t *d = c->d;
    const size_t psize = (size_t) d->meqn * (d->mx + 2 * d->mbc) * (d->my + 2 * d->mbc);
    int p;
    for (p = 0; p < d->npatch; p++) {
        if (d->level[p] != level) continue;
        oracle_timeinterp_patch(d->mx, d->my, d->mbc, d->meqn, PATCH(c->q, p), PATCH(c->qlast, p),
                                PATCH(c->qti, p), alpha);
    }
}

static double advance_level_rec(sub_ctx_t *c, int level, int curr_fine_step, double maxcfl)
{
    const oracle_domain_t *d = c->d;
    int p;
    for (p = 0; p < d->npatch; p++)
        if (d->level[p] == level)
            maxcfl = fmax(maxcfl, update_patch(d, c->claw, c->q, c->aux, c->qlast, c->reg, c->fl,
                                               p, c->dt_step[level]));
    c->last_step[level] += c->step_inc[level];
    c->current_time[level] += c->dt_step[level];
    if (level > c->minlevel) {
        const int coarser = level - 1;
        if (c->last_step[coarser] == curr_fine_step) {
            maxcfl = fmax(maxcfl, advance_level_rec(c, coarser, c->last_step[coarser], maxcfl));
            /* compute_alpha (src/fclaw2d_advance.c:152-170) */
            timeinterp_level(c, coarser, ((double) (c->last_step[level] % c->step_inc[coarser])) /
                                             c->step_inc[coarser]);
        }
    }
    return maxcfl;
}

double oracle_advance_subcycle(const oracle_domain_t *d, const oracle_claw_t *claw, double *q,
                               const double *aux, double dt, int time_sync)
{
    const int mx = d->mx, my = d->my, mbc = d->mbc, meqn = d->meqn;
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    sub_ctx_t c;
    double maxcfl = 0.0, dtl;
    int level, nf, inc, tot;
    if (d->npatch <= 0) return 0.0;
    memset(&c, 0, sizeof(c));
    c.d = d;
    c.claw = claw;
    c.q = q;
    c.aux = aux;
    domain_levels(d, &c.h, &c.minlevel, &c.maxlevel);
    if (c.maxlevel >= 31) { hash_free(&c.h); return -1.0; }
    c.qlast = (double *) calloc((size_t) d->npatch * psize, sizeof(double));
    c.qti = (double *) calloc((size_t) d->npatch * psize, sizeof(double));
    if (time_sync) {
        c.fl = (double *) malloc(4 * psize * sizeof(double));
        c.reg = (double *) calloc((size_t) d->npatch * oracle_sync_register_count(mx, meqn),
                                  sizeof(double));
    }
    /* initialize_timestep_counters (:54-110) with subcycle */
    dtl = dt;
    inc = 1 << (c.maxlevel - c.minlevel);
    tot = 1;
    for (level = c.minlevel; level <= c.maxlevel; level++) {
        c.last_step[level] = 0;
        c.dt_step[level] = dtl;
        c.total_steps[level] = tot;
        c.step_inc[level] = inc;
        dtl /= 2;
        inc /= 2;
        tot *= 2;
    }
    for (nf = 0; nf < c.total_steps[c.maxlevel]; nf++) {
        int last_step;
        maxcfl = fmax(maxcfl, advance_level_rec(&c, c.maxlevel, nf, maxcfl));
        last_step = c.last_step[c.maxlevel];
        if (last_step < c.total_steps[c.maxlevel]) {
            int ti = c.maxlevel - 1;
            while (last_step % c.step_inc[ti] == 0) ti--;
            ghost_update_range(d, &c.h, q, c.qti, ti + 1, c.maxlevel, 1);
            if (time_sync) time_sync_range(d, &c.h, q, c.reg, ti + 1, c.maxlevel);
        }
    }
    ghost_update_range(d, &c.h, q, NULL, c.minlevel, c.maxlevel, 0);
    if (time_sync) time_sync_range(d, &c.h, q, c.reg, c.minlevel, c.maxlevel);
    free(c.qlast);
    free(c.qti);
    free(c.fl);
    free(c.reg);
    hash_free(&c.h);
    return maxcfl;
}

void oracle_mass(const oracle_domain_t *d, const double *q, double *sum)
{
    const int mx = d->mx, my = d->my, mbc = d->mbc, meqn = d->meqn;
    const size_t psize = (size_t) meqn * (mx + 2 * mbc) * (my + 2 * mbc);
    const double wx = (d->bx - d->ax) / d->nbx, wy = (d->by - d->ay) / d->nby;
    double *c = (double *) calloc((size_t) meqn, sizeof(double));
    int p, i, j, m;
    for (m = 0; m < meqn; m++) sum[m] = 0.0;
    for (p = 0; p < d->npatch; p++) {
        const double nl = (double) (1 << d->level[p]);
        const double dx = wx / (nl * mx), dy = wy / (nl * my);
        const double area = dx * dy;
        const double *qp = PATCH(q, p);
        for (m = 0; m < meqn; m++)
            for (j = 1; j <= my; j++)
                for (i = 1; i <= mx; i++) {
                    double y = Q(qp, i, j, m) * area - c[m];
                    double t = sum[m] + y;
                    c[m] = (t - sum[m]) - y;
                    sum[m] = t;
                }
    }
    free(c);
}
