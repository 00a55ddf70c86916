/*
 * Host forest and ghost-fill program builder (declared in
 * include/fclaw2d_hip_forest.h).
 *
 * Neighbour relations follow p4est's semantics as ForestClaw uses them:
 *   faces   : same size, half size (two neighbours) or double size
 *             (src/fclaw2d_face_neighbors.c:63-190)
 *   corners : a corner neighbour exists unless the corner is a hanging node,
 *             i.e. lies inside a coarser patch's face
 *             (src/fclaw2d_corner_neighbors.c:95-330)
 * Every block of the brick has the same orientation, so all index transforms
 * (src/forestclaw2d.c:608-977, 1128-1380) reduce to translations.
 */
#include "../../include/fclaw2d_hip_forest.h"

#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { REL_BOUNDARY = 0, REL_SAME = 1, REL_HALF = 2, REL_DOUBLE = 3, REL_NONE = 4 };

typedef struct {
    int rel;
    int nb[2];
    int ti[2], tj[2];
} nbinfo_t;

/* One slot of the leaf map: key and leaf side by side, so a probe touches
 * one cache line.  An empty slot has val < 0 (key 0 is a valid key). */
typedef struct {
    uint64_t key;
    int64_t val;
} hentry_t;

struct fclaw2d_hip_forest {
    int nbx, nby, per_x, per_y;
    int iso;   /* blocks are separate domains (an ensemble of runs) */
    int bal;   /* the leaves are known to be 2:1 balanced */
    int n, cap;
    int *level, *gx, *gy;
    /* provenance of the leaves after the last adapt: kind and 4 sources each */
    int *pkind, *psrc, pn;
    /* open-addressing map (level, gx, gy) -> leaf; hok: it holds the current
     * leaves (kept up to date through splits, coarsenings and sorts, rebuilt
     * only when it would fill beyond half) */
    size_t hcap;
    hentry_t *hash;
    int hok;
    /* face (4) and corner (4) relations of every leaf for one (mx, my), built
     * on first use and dropped whenever the leaves change */
    nbinfo_t *rel;
    int rel_mx, rel_my, rel_valid;
    size_t rel_cap;          /* leaves rel has room for */
    nbinfo_t *rel_spare;     /* a second buffer, for carrying relations over an adapt */
    size_t spare_cap;
};

/* ---------------------------------------------------------------- hashing */

static uint64_t fkey(int level, int64_t gx, int64_t gy)
{
    return ((uint64_t) level << 58) | ((uint64_t) gx << 29) | (uint64_t) gy;
}

static size_t fslot(uint64_t k, size_t cap)
{
    return (size_t) ((k * 0x9E3779B97F4A7C15ull) >> 17) & (cap - 1);
}

static int rehash(fclaw2d_hip_forest_t *f)
{
    size_t cap = 64, p;
    int i;
    f->rel_valid = 0;
    f->hok = 0;
    while (cap < (size_t) f->n * 2) cap <<= 1;   /* load factor <= 1/2 */
    if (cap != f->hcap) {
        free(f->hash);
        f->hash = (hentry_t *) malloc(cap * sizeof(hentry_t));
        if (!f->hash) { f->hcap = 0; return -1; }
        f->hcap = cap;
    }
    for (p = 0; p < cap; p++) f->hash[p].val = -1;
    for (i = 0; i < f->n; i++) {
        uint64_t k = fkey(f->level[i], f->gx[i], f->gy[i]);
        p = fslot(k, cap);
        while (f->hash[p].val >= 0) {
            if (f->hash[p].key == k) return -1; /* duplicate leaf */
            p = (p + 1) & (cap - 1);
        }
        f->hash[p].key = k;
        f->hash[p].val = i;
    }
    f->hok = 1;
    return 0;
}

/* Add leaf v with key k to the map (-1: k is there already). */
static int hash_put(fclaw2d_hip_forest_t *f, uint64_t k, int v)
{
    size_t p = fslot(k, f->hcap);
    while (f->hash[p].val >= 0) {
        if (f->hash[p].key == k) return -1;
        p = (p + 1) & (f->hcap - 1);
    }
    f->hash[p].key = k;
    f->hash[p].val = v;
    return 0;
}

/* Remove key k from the map (-1: not there).  Linear probing without
 * tombstones: the entries after the hole in its run move back into it when
 * their home slot does not lie cyclically after the hole. */
static int hash_del(fclaw2d_hip_forest_t *f, uint64_t k)
{
    const size_t mask = f->hcap - 1;
    size_t p = fslot(k, f->hcap), q;
    while (f->hash[p].val >= 0 && f->hash[p].key != k) p = (p + 1) & mask;
    if (f->hash[p].val < 0) return -1;
    for (q = (p + 1) & mask; f->hash[q].val >= 0; q = (q + 1) & mask) {
        const size_t h = fslot(f->hash[q].key, f->hcap);
        if (q > p ? (h <= p || h > q) : (h <= p && h > q)) {
            f->hash[p] = f->hash[q];
            p = q;
        }
    }
    f->hash[p].val = -1;
    return 0;
}

/* Renumber the map's leaves: leaf i becomes perm[i] (perm[i] >= 0). */
static void hash_renumber(fclaw2d_hip_forest_t *f, const int *perm)
{
    size_t p;
    for (p = 0; p < f->hcap; p++)
        if (f->hash[p].val >= 0) f->hash[p].val = perm[f->hash[p].val];
}

static int lookup(const fclaw2d_hip_forest_t *f, int level, int64_t gx, int64_t gy)
{
    uint64_t k;
    size_t p;
    if (level < 0 || gx < 0 || gy < 0) return -1;
    k = fkey(level, gx, gy);
    p = fslot(k, f->hcap);
    while (f->hash[p].val >= 0) {
        if (f->hash[p].key == k) return (int) f->hash[p].val;
        p = (p + 1) & (f->hcap - 1);
    }
    return -1;
}

/* Map a level-L patch coordinate into the brick; 0 if it falls outside a
 * non-periodic side.  (sx, sy) is a level-L coordinate in the block the
 * query starts from: when the blocks are separate domains (an ensemble of
 * runs), a coordinate outside that block is outside the domain, or wraps
 * around that block alone when periodic. */
static int wrap(const fclaw2d_hip_forest_t *f, int L, int64_t *x, int64_t *y, int64_t sx, int64_t sy)
{
    const int64_t nx = (int64_t) f->nbx << L, ny = (int64_t) f->nby << L;
    if (f->iso) {
        const int64_t w = (int64_t) 1 << L, bx0 = (sx >> L) * w, by0 = (sy >> L) * w;
        int64_t lx = *x - bx0, ly = *y - by0;
        if (lx < 0 || lx >= w) {
            if (!f->per_x) return 0;
            lx = ((lx % w) + w) % w;
        }
        if (ly < 0 || ly >= w) {
            if (!f->per_y) return 0;
            ly = ((ly % w) + w) % w;
        }
        *x = bx0 + lx;
        *y = by0 + ly;
        return 1;
    }
    if (*x < 0 || *x >= nx) {
        if (!f->per_x) return 0;
        *x = ((*x % nx) + nx) % nx;
    }
    if (*y < 0 || *y >= ny) {
        if (!f->per_y) return 0;
        *y = ((*y % ny) + ny) % ny;
    }
    return 1;
}

/* Leaf covering level-L cell (x, y) at level <= L, or -1 if that region is
 * refined beyond L (or outside the domain). */
static int covering(const fclaw2d_hip_forest_t *f, int L, int64_t x, int64_t y, int64_t sx,
                    int64_t sy)
{
    int l;
    if (!wrap(f, L, &x, &y, sx, sy)) return -1;
    for (l = L; l >= 0; l--) {
        int q = lookup(f, l, x >> (L - l), y >> (L - l));
        if (q >= 0) return q;
    }
    return -1;
}

/* ---------------------------------------------------------------- ordering */

static uint64_t spread(uint32_t v)
{
    uint64_t x = v;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}

/* One stable counting pass of an LSD radix sort: order idx by byte `shift`
 * of key[idx]. */
static void radix_pass(const uint64_t *key, const int *in, int *out, int n, int shift)
{
    int count[257] = {0}, i;
    for (i = 0; i < n; i++) count[((key[in[i]] >> shift) & 255) + 1]++;
    for (i = 0; i < 256; i++) count[i + 1] += count[i];
    for (i = 0; i < n; i++) out[count[(key[in[i]] >> shift) & 255]++] = in[i];
}

/* p4est order: by block (row-major), then the Morton index of the leaf's
 * lower-left corner at the finest level, then level (a parent's corner is its
 * first child's).  LSD radix sort over (level, Morton, block), as stable
 * counting passes over the bytes that vary. */
static int sort_leaves(fclaw2d_hip_forest_t *f)
{
    const int n = f->n;
    uint64_t *mort = (uint64_t *) malloc(sizeof(uint64_t) * (size_t) (n + 1));
    uint64_t *lvbk = (uint64_t *) malloc(sizeof(uint64_t) * (size_t) (n + 1));
    int *a = (int *) malloc(sizeof(int) * (size_t) (n + 1));
    int *b = (int *) malloc(sizeof(int) * (size_t) (n + 1));
    int *tmp = (int *) malloc(sizeof(int) * 3 * (size_t) (n + 1));
    uint64_t mor_or = 0, bk_or = 0;
    int i, sh;
    if (!mort || !lvbk || !a || !b || !tmp) {
        free(mort); free(lvbk); free(a); free(b); free(tmp);
        return -1;
    }
    for (i = 0; i < n; i++) {
        const int L = f->level[i];
        const uint32_t bx = (uint32_t) (f->gx[i] >> L), by = (uint32_t) (f->gy[i] >> L);
        const uint32_t lx = (uint32_t) (f->gx[i] - ((int64_t) bx << L));
        const uint32_t ly = (uint32_t) (f->gy[i] - ((int64_t) by << L));
        const uint64_t block = (uint64_t) by * (uint64_t) f->nbx + bx;
        mort[i] = spread(lx << (28 - L)) | (spread(ly << (28 - L)) << 1);
        lvbk[i] = (uint64_t) L | (block << 8);
        mor_or |= mort[i];
        bk_or |= block;
        a[i] = i;
    }
    radix_pass(lvbk, a, b, n, 0);                       /* level (< 256) */
    { int *t = a; a = b; b = t; }
    for (sh = 0; sh < 64; sh += 8)                      /* Morton bytes in use */
        if ((mor_or >> sh) & 255) { radix_pass(mort, a, b, n, sh); int *t = a; a = b; b = t; }
    for (sh = 8; sh < 64; sh += 8)                      /* block bytes in use */
        if ((bk_or >> (sh - 8)) & 255) { radix_pass(lvbk, a, b, n, sh); int *t = a; a = b; b = t; }
    for (i = 0; i < n; i++) {
        tmp[3 * i] = f->level[a[i]]; tmp[3 * i + 1] = f->gx[a[i]]; tmp[3 * i + 2] = f->gy[a[i]];
    }
    for (i = 0; i < n; i++) { f->level[i] = tmp[3 * i]; f->gx[i] = tmp[3 * i + 1]; f->gy[i] = tmp[3 * i + 2]; }
    f->rel_valid = 0;
    if (f->hok) {   /* the map holds these leaves: renumber it (b: old -> new) */
        for (i = 0; i < n; i++) b[a[i]] = i;
        hash_renumber(f, b);
    }
    free(mort); free(lvbk); free(a); free(b); free(tmp);
    return f->hok ? 0 : rehash(f);
}

static int reserve(fclaw2d_hip_forest_t *f, int n)
{
    if (n <= f->cap) return 0;
    {
        int cap = f->cap ? f->cap : 64;
        while (cap < n) cap *= 2;
        int *l = (int *) realloc(f->level, (size_t) cap * sizeof(int));
        if (!l) return -1;
        f->level = l;
        l = (int *) realloc(f->gx, (size_t) cap * sizeof(int));
        if (!l) return -1;
        f->gx = l;
        l = (int *) realloc(f->gy, (size_t) cap * sizeof(int));
        if (!l) return -1;
        f->gy = l;
        f->cap = cap;
    }
    return 0;
}

/* Replace every marked leaf by its four children (the first child keeps the
 * leaf's index, the others are appended). */
static int split_marked(fclaw2d_hip_forest_t *f, const char *mark)
{
    int i, add = 0, n0 = f->n, keep;
    for (i = 0; i < n0; i++) add += mark[i] ? 3 : 0;
    if (reserve(f, n0 + add)) return -1;
    keep = f->hok && (size_t) (n0 + add) * 2 <= f->hcap;   /* the map stays at most half full */
    f->hok = 0;
    f->rel_valid = 0;
    for (i = 0; i < n0; i++) {
        if (!mark[i]) continue;
        const int L = f->level[i] + 1, x = 2 * f->gx[i], y = 2 * f->gy[i];
        if (keep && hash_del(f, fkey(L - 1, f->gx[i], f->gy[i]))) return -1;
        f->level[i] = L; f->gx[i] = x; f->gy[i] = y;
        f->level[f->n] = L; f->gx[f->n] = x + 1; f->gy[f->n] = y;     f->n++;
        f->level[f->n] = L; f->gx[f->n] = x;     f->gy[f->n] = y + 1; f->n++;
        f->level[f->n] = L; f->gx[f->n] = x + 1; f->gy[f->n] = y + 1; f->n++;
        if (keep && (hash_put(f, fkey(L, x, y), i) || hash_put(f, fkey(L, x + 1, y), f->n - 3) ||
                     hash_put(f, fkey(L, x, y + 1), f->n - 2) || hash_put(f, fkey(L, x + 1, y + 1), f->n - 1)))
            return -1;
    }
    f->hok = keep;
    return 0;
}

/* Restore the 2:1 face and corner balance by splitting leaves that are more
 * than one level coarser than a neighbour.  The first pass checks the leaves
 * of work0 (all leaves when work0 is null); later passes check only the
 * leaves a split created and the leaves that asked for it (a leaf no pass
 * revisits saw no coarser neighbour, and splits only make its neighbours
 * finer).  On a balanced forest whose leaves work0 were just split, only
 * their children can see a neighbour two levels coarser, so work0 = the
 * children suffices. */
static int balance_from(fclaw2d_hip_forest_t *f, const int *work0, int nwork0)
{
    static const int dx8[8] = {-1, 1, 0, 0, -1, 1, -1, 1};
    static const int dy8[8] = {0, 0, -1, 1, -1, -1, 1, 1};
    int *work = NULL, nwork = work0 ? nwork0 : f->n, first = 1, rc = 0;
    for (;;) {
        char *mark, *trig;
        int *next, nnext = 0, w, d, any = 0, n0;
        if (!f->hok && rehash(f)) { rc = -1; break; }
        mark = (char *) calloc((size_t) f->n + 1, 1);
        trig = (char *) calloc((size_t) f->n + 1, 1);
        if (!mark || !trig) { free(mark); free(trig); rc = -1; break; }
        for (w = 0; w < nwork; w++) {
            const int i = first ? (work0 ? work0[w] : w) : work[w];
            const int L = f->level[i];
            if (L < 2) continue;
            for (d = 0; d < 8; d++) {
                const int q = covering(f, L, (int64_t) f->gx[i] + dx8[d], (int64_t) f->gy[i] + dy8[d],
                                       f->gx[i], f->gy[i]);
                if (q >= 0 && f->level[q] < L - 1) {
                    trig[i] = 1;
                    if (!mark[q]) { mark[q] = 1; any = 1; }
                }
            }
        }
        if (!any) { free(mark); free(trig); break; }
        n0 = f->n;
        if (split_marked(f, mark)) { free(mark); free(trig); rc = -1; break; }
        /* next pass: the split leaves (their first child kept the index),
         * the appended children and the leaves that triggered a split */
        next = (int *) malloc(sizeof(int) * (size_t) (f->n + n0 + 1));
        if (!next) { free(mark); free(trig); rc = -1; break; }
        for (w = 0; w < n0; w++)
            if (mark[w] || trig[w]) next[nnext++] = w;
        for (w = n0; w < f->n; w++) next[nnext++] = w;
        free(mark); free(trig);
        free(work);
        work = next; nwork = nnext; first = 0;
    }
    free(work);
    if (rc) return rc;
    f->bal = 1;
    return sort_leaves(f);
}

static int balance(fclaw2d_hip_forest_t *f) { return balance_from(f, NULL, 0); }

/* Split the marked leaves and restore the balance, checking only the new
 * children when the forest was balanced before. */
static int split_and_balance(fclaw2d_hip_forest_t *f, const char *mark)
{
    const int n0 = f->n;
    int *kids, nk = 0, i, rc, any = 0;
    for (i = 0; i < n0 && !any; i++) any = mark[i] != 0;
    if (!any && f->bal) return 0;   /* nothing split: the leaves, their order and map stay */
    if (split_marked(f, mark)) return -1;
    if (!f->bal) return balance(f);
    kids = (int *) malloc(sizeof(int) * (size_t) (f->n + 1));
    if (!kids) return -1;
    for (i = 0; i < n0; i++)
        if (mark[i]) kids[nk++] = i;           /* the first child keeps the parent's index */
    for (i = n0; i < f->n; i++) kids[nk++] = i;
    rc = balance_from(f, kids, nk);
    free(kids);
    return rc;
}

/* ---------------------------------------------------------------- public */

static fclaw2d_hip_forest_t *alloc_forest(int nbx, int nby, int px, int py)
{
    fclaw2d_hip_forest_t *f;
    if (nbx <= 0 || nby <= 0) return NULL;
    f = (fclaw2d_hip_forest_t *) calloc(1, sizeof(*f));
    if (!f) return NULL;
    f->nbx = nbx; f->nby = nby; f->per_x = px != 0; f->per_y = py != 0;
    return f;
}

fclaw2d_hip_forest_t *fclaw2d_hip_forest_new(int nbx, int nby, int periodic_x, int periodic_y,
                                             int minlevel)
{
    fclaw2d_hip_forest_t *f;
    int x, y;
    if (minlevel < 0 || minlevel > 20) return NULL;
    f = alloc_forest(nbx, nby, periodic_x, periodic_y);
    if (!f) return NULL;
    if (reserve(f, (nbx * nby) << (2 * minlevel))) { fclaw2d_hip_forest_destroy(f); return NULL; }
    for (y = 0; y < (nby << minlevel); y++)
        for (x = 0; x < (nbx << minlevel); x++) {
            f->level[f->n] = minlevel; f->gx[f->n] = x; f->gy[f->n] = y; f->n++;
        }
    f->bal = 1;   /* one level */
    if (sort_leaves(f)) { fclaw2d_hip_forest_destroy(f); return NULL; }
    return f;
}

fclaw2d_hip_forest_t *fclaw2d_hip_forest_from_leaves(int nbx, int nby, int periodic_x,
                                                     int periodic_y, int n, const int *level,
                                                     const int *gx, const int *gy, int bal)
{
    fclaw2d_hip_forest_t *f = alloc_forest(nbx, nby, periodic_x, periodic_y);
    double area = 0.0;
    int i;
    if (!f || n <= 0 || reserve(f, n)) { fclaw2d_hip_forest_destroy(f); return NULL; }
    for (i = 0; i < n; i++) {
        if (level[i] < 0 || level[i] > 28 || gx[i] < 0 || gy[i] < 0 ||
            gx[i] >= (nbx << level[i]) || gy[i] >= (nby << level[i])) {
            fclaw2d_hip_forest_destroy(f);
            return NULL;
        }
        f->level[i] = level[i]; f->gx[i] = gx[i]; f->gy[i] = gy[i];
        area += 1.0 / (double) (1ull << (2 * level[i]));
    }
    f->n = n;
    /* leaves must be distinct and cover the brick exactly */
    if (area != (double) (nbx * nby) || rehash(f)) { fclaw2d_hip_forest_destroy(f); return NULL; }
    for (i = 0; i < n; i++) {
        int l;
        for (l = level[i] - 1; l >= 0; l--)
            if (lookup(f, l, gx[i] >> (level[i] - l), gy[i] >> (level[i] - l)) >= 0) {
                fclaw2d_hip_forest_destroy(f);
                return NULL;
            }
    }
    if ((bal ? balance(f) : sort_leaves(f)) != 0) { fclaw2d_hip_forest_destroy(f); return NULL; }
    return f;
}

int fclaw2d_hip_forest_set_isolated(fclaw2d_hip_forest_t *f, int isolated)
{
    if (!f) return -1;
    f->iso = isolated != 0;
    f->rel_valid = 0;
    return 0;
}

void fclaw2d_hip_forest_destroy(fclaw2d_hip_forest_t *f)
{
    if (!f) return;
    free(f->level); free(f->gx); free(f->gy); free(f->hash);
    free(f->pkind); free(f->psrc); free(f->rel); free(f->rel_spare);
    free(f);
}

int fclaw2d_hip_forest_num_patches(const fclaw2d_hip_forest_t *f) { return f ? f->n : 0; }

int fclaw2d_hip_forest_min_level(const fclaw2d_hip_forest_t *f)
{
    int i, m = 1 << 30;
    for (i = 0; i < f->n; i++) if (f->level[i] < m) m = f->level[i];
    return m;
}

int fclaw2d_hip_forest_max_level(const fclaw2d_hip_forest_t *f)
{
    int i, m = -1;
    for (i = 0; i < f->n; i++) if (f->level[i] > m) m = f->level[i];
    return m;
}

void fclaw2d_hip_forest_leaves(const fclaw2d_hip_forest_t *f, int *level, int *gx, int *gy)
{
    memcpy(level, f->level, (size_t) f->n * sizeof(int));
    memcpy(gx, f->gx, (size_t) f->n * sizeof(int));
    memcpy(gy, f->gy, (size_t) f->n * sizeof(int));
}

int fclaw2d_hip_forest_refine(fclaw2d_hip_forest_t *f, const int *flags, int maxlevel)
{
    char *mark;
    int i, rc;
    if (!f || !flags) return -1;
    f->pn = -1;  /* provenance describes adapt only */
    mark = (char *) calloc((size_t) f->n, 1);
    if (!mark) return -1;
    for (i = 0; i < f->n; i++) mark[i] = flags[i] != 0 && f->level[i] < maxlevel;
    rc = split_and_balance(f, mark);
    free(mark);
    return rc ? -1 : f->n;
}

/* ---------------------------------------------------------------- adapt */

/* Can the four children of level-(L-1) patch (X, Y) be replaced by their
 * parent without breaking the 2:1 face and corner balance?  Every level-L
 * cell in the ring around the parent must be covered by a leaf of level <= L. */
static int coarsen_keeps_balance(const fclaw2d_hip_forest_t *f, int L, int64_t X, int64_t Y)
{
    int64_t x, y;
    for (y = 2 * Y - 1; y <= 2 * Y + 2; y++)
        for (x = 2 * X - 1; x <= 2 * X + 2; x++) {
            int64_t wx = x, wy = y;
            if (x >= 2 * X && x <= 2 * X + 1 && y >= 2 * Y && y <= 2 * Y + 1) continue;
            if (!wrap(f, L, &wx, &wy, 2 * X, 2 * Y)) continue;          /* outside the domain */
            if (covering(f, L, wx, wy, 2 * X, 2 * Y) < 0) return 0;     /* refined beyond L */
        }
    return 1;
}

static void face_rel(const fclaw2d_hip_forest_t *f, int p, int iface, int mx, int my, nbinfo_t *r);
static void corner_rel(const fclaw2d_hip_forest_t *f, int p, int ic, int mx, int my, nbinfo_t *r);

/* The relation cache of the adapted leaves from the one of the leaves before
 * the adapt: a leaf that stayed (provenance SAME) keeps each relation whose
 * neighbours all stayed, with their new numbers (the neighbour still covers
 * the same region, so the kind and the index translation are unchanged);
 * every other relation, and every relation of a new leaf, is recomputed.  A
 * corner without a neighbour (a hanging node) names no leaf, so it is always
 * recomputed. */
static int update_relations(fclaw2d_hip_forest_t *f, const nbinfo_t *orel, int n0, int mx, int my)
{
    int *o2n = (int *) malloc(sizeof(int) * (size_t) (n0 + 1));
    nbinfo_t *rel;
    int i, k, m;
    if (!o2n) return -1;
    if (f->spare_cap < (size_t) f->n) {   /* grown with headroom, kept across adapts */
        free(f->rel_spare);
        f->spare_cap = (size_t) f->n + (size_t) f->n / 4 + 16;
        f->rel_spare = (nbinfo_t *) malloc(sizeof(nbinfo_t) * 8 * f->spare_cap);
        if (!f->rel_spare) { f->spare_cap = 0; free(o2n); return -1; }
    }
    rel = f->rel_spare;
    for (i = 0; i < n0; i++) o2n[i] = -1;
    for (i = 0; i < f->n; i++)
        if (f->pkind[i] == FCLAW2D_HIP_REGRID_SAME) o2n[f->psrc[4 * i]] = i;
    for (i = 0; i < f->n; i++) {
        const int q = f->pkind[i] == FCLAW2D_HIP_REGRID_SAME ? f->psrc[4 * i] : -1;
        int kept[4] = {0, 0, 0, 0};
        for (k = 0; k < 8; k++) {
            nbinfo_t *r = &rel[8 * (size_t) i + k];
            int keep = q >= 0;
            if (keep) {
                *r = orel[8 * (size_t) q + k];
                if (r->rel == REL_NONE) {
                    /* a hanging corner lies inside the side of a coarser leaf,
                     * which is the face neighbour on one of the corner's two
                     * faces: while both face relations stand, so does it */
                    const int ic = k - 4;
                    keep = k >= 4 && kept[(ic & 1) ? 1 : 0] && kept[(ic & 2) ? 3 : 2];
                } else if (r->rel != REL_BOUNDARY) {
                    const int nn = (r->rel == REL_HALF && k < 4) ? 2 : 1;
                    for (m = 0; m < nn && keep; m++) {
                        r->nb[m] = o2n[r->nb[m]];
                        keep = r->nb[m] >= 0;
                    }
                }
            }
            if (!keep) {
                if (k < 4) face_rel(f, i, k, mx, my, r);
                else corner_rel(f, i, k - 4, mx, my, r);
            }
            if (k < 4) kept[k] = keep;
        }
    }
    free(o2n);
    free(f->rel);                 /* (null: the adapt holds the old buffer) */
    f->rel = rel;
    f->rel_cap = f->spare_cap;
    f->rel_spare = NULL;
    f->spare_cap = 0;
    f->rel_mx = mx; f->rel_my = my; f->rel_valid = 1;
    return 0;
}

int fclaw2d_hip_forest_adapt(fclaw2d_hip_forest_t *f, const int *refine, const int *coarsen,
                             int minlevel, int maxlevel)
{
    fclaw2d_hip_forest_t *old;
    char *mark;
    int i, g, n0, ncoarse = 0;
    int *fam = NULL;   /* parent (level, X, Y) of families to coarsen, 3 ints each */
    nbinfo_t *orel = NULL;   /* the relation cache of the leaves before the adapt */
    size_t orel_cap = 0;
    int omx = 0, omy = 0;
    if (!f || !refine || !coarsen) return -1;
    n0 = f->n;
    if (f->rel_valid) {
        orel = f->rel;
        orel_cap = f->rel_cap;
        omx = f->rel_mx; omy = f->rel_my;
        f->rel = NULL;
        f->rel_cap = 0;
        f->rel_valid = 0;
    }
    /* keep the old leaves for the provenance (with a copy of their map) */
    old = alloc_forest(f->nbx, f->nby, f->per_x, f->per_y);
    if (!old || reserve(old, n0)) { fclaw2d_hip_forest_destroy(old); free(orel); return -1; }
    memcpy(old->level, f->level, (size_t) n0 * sizeof(int));
    memcpy(old->gx, f->gx, (size_t) n0 * sizeof(int));
    memcpy(old->gy, f->gy, (size_t) n0 * sizeof(int));
    old->n = n0;
    if (f->hok) {
        old->hash = (hentry_t *) malloc(f->hcap * sizeof(hentry_t));
        if (!old->hash) { fclaw2d_hip_forest_destroy(old); free(orel); return -1; }
        memcpy(old->hash, f->hash, f->hcap * sizeof(hentry_t));
        old->hcap = f->hcap;
        old->hok = 1;
    } else if (rehash(old) || rehash(f)) {
        fclaw2d_hip_forest_destroy(old); free(orel); return -1;
    }
    /* families to coarsen: all four siblings flagged, none flagged to refine
     * (cb_regrid_tag4coarsening, src/fclaw2d_regrid.c:61-93) */
    fam = (int *) malloc(sizeof(int) * 3 * (size_t) (n0 / 4 + 1));
    if (!fam) { fclaw2d_hip_forest_destroy(old); free(orel); return -1; }
    for (i = 0; i < n0; i++) {
        const int L = f->level[i], x = f->gx[i], y = f->gy[i];
        int sib[4], ok = L > minlevel && L > 0 && !(x & 1) && !(y & 1);
        for (g = 0; g < 4 && ok; g++) {
            sib[g] = lookup(old, L, x + (g & 1), y + (g >> 1));
            ok = sib[g] >= 0 && coarsen[sib[g]] && !(refine[sib[g]] && L < maxlevel);
        }
        if (ok) { fam[3 * ncoarse] = L; fam[3 * ncoarse + 1] = x >> 1; fam[3 * ncoarse + 2] = y >> 1; ncoarse++; }
    }
    /* refine (cb_fclaw2d_regrid_tag4refinement) and restore the balance */
    mark = (char *) calloc((size_t) n0 + 1, 1);
    if (!mark) { free(fam); fclaw2d_hip_forest_destroy(old); free(orel); return -1; }
    for (i = 0; i < n0; i++) mark[i] = refine[i] != 0 && f->level[i] < maxlevel;
    if (split_and_balance(f, mark)) { free(mark); free(fam); fclaw2d_hip_forest_destroy(old); free(orel); return -1; }
    free(mark);
    /* coarsen the families that are still intact and keep the balance */
    if (ncoarse > 0) {
        int k, w, keep = f->hok;
        char *drop = (char *) calloc((size_t) f->n + 1, 1);
        int *cmap = (int *) malloc(sizeof(int) * ((size_t) f->n + 1));
        int nadd = 0;
        if (!drop || !cmap) { free(drop); free(cmap); free(fam); fclaw2d_hip_forest_destroy(old); free(orel); return -1; }
        for (k = 0; k < ncoarse; k++) {
            const int L = fam[3 * k], X = fam[3 * k + 1], Y = fam[3 * k + 2];
            int kids[4], ok = 1;
            for (g = 0; g < 4 && ok; g++) {
                kids[g] = lookup(f, L, 2 * X + (g & 1), 2 * Y + (g >> 1));
                ok = kids[g] >= 0;
            }
            if (!ok || !coarsen_keeps_balance(f, L, X, Y)) { fam[3 * k] = -1; continue; }
            for (g = 0; g < 4; g++) drop[kids[g]] = 1;
            nadd++;
        }
        /* the map: the children leave, the kept leaves are renumbered by the
         * compaction, the parents join */
        f->hok = 0;
        for (i = 0; i < f->n && keep; i++)
            if (drop[i]) keep = hash_del(f, fkey(f->level[i], f->gx[i], f->gy[i])) == 0;
        for (i = 0, w = 0; i < f->n; i++)
            if (!drop[i]) { cmap[i] = w; f->level[w] = f->level[i]; f->gx[w] = f->gx[i]; f->gy[w] = f->gy[i]; w++; }
        if (keep) hash_renumber(f, cmap);
        f->n = w;
        free(drop);
        free(cmap);
        if (reserve(f, f->n + nadd)) { free(fam); fclaw2d_hip_forest_destroy(old); free(orel); return -1; }
        for (k = 0; k < ncoarse; k++)
            if (fam[3 * k] >= 0) {
                f->level[f->n] = fam[3 * k] - 1; f->gx[f->n] = fam[3 * k + 1]; f->gy[f->n] = fam[3 * k + 2];
                if (keep) keep = hash_put(f, fkey(f->level[f->n], f->gx[f->n], f->gy[f->n]), f->n) == 0;
                f->n++;
            }
        f->hok = keep;
        f->rel_valid = 0;
        if (sort_leaves(f)) { free(fam); fclaw2d_hip_forest_destroy(old); free(orel); return -1; }
    }
    free(fam);
    /* provenance of every new leaf (cb_fclaw2d_regrid_repopulate, src/fclaw2d_regrid.c:95-230) */
    free(f->pkind); free(f->psrc);
    f->pkind = (int *) malloc(sizeof(int) * (size_t) f->n);
    f->psrc = (int *) malloc(sizeof(int) * 4 * (size_t) f->n);
    if (!f->pkind || !f->psrc) { fclaw2d_hip_forest_destroy(old); free(orel); return -1; }
    f->pn = f->n;
    for (i = 0; i < f->n; i++) {
        const int L = f->level[i], x = f->gx[i], y = f->gy[i];
        int q = lookup(old, L, x, y);
        int *src = f->psrc + 4 * i;
        src[0] = src[1] = src[2] = src[3] = -1;
        if (q >= 0) { f->pkind[i] = FCLAW2D_HIP_REGRID_SAME; src[0] = q; continue; }
        if (L > 0 && (q = lookup(old, L - 1, x >> 1, y >> 1)) >= 0) {
            f->pkind[i] = FCLAW2D_HIP_REGRID_CHILD; src[0] = q; src[1] = (x & 1) + 2 * (y & 1);
            continue;
        }
        f->pkind[i] = FCLAW2D_HIP_REGRID_PARENT;
        for (g = 0; g < 4; g++) {
            src[g] = lookup(old, L + 1, 2 * x + (g & 1), 2 * y + (g >> 1));
            if (src[g] < 0) { fclaw2d_hip_forest_destroy(old); free(orel); return -1; }
        }
    }
    fclaw2d_hip_forest_destroy(old);
    if (orel) {
        const int rc = update_relations(f, orel, n0, omx, omy);
        if (rc) { free(orel); return -1; }
        /* the old buffer is the spare of the next adapt */
        free(f->rel_spare);
        f->rel_spare = orel;
        f->spare_cap = orel_cap;
    }
    return f->n;
}

int fclaw2d_hip_forest_provenance(const fclaw2d_hip_forest_t *f, int *kind, int *src)
{
    if (!f || !f->pkind || f->pn != f->n) return -1;
    memcpy(kind, f->pkind, sizeof(int) * (size_t) f->n);
    memcpy(src, f->psrc, sizeof(int) * 4 * (size_t) f->n);
    return f->n;
}

/* ---------------------------------------------------------------- neighbours */

static const int fdx[4] = {-1, 1, 0, 0}, fdy[4] = {0, 0, -1, 1};
static const int cdx[4] = {-1, 1, -1, 1}, cdy[4] = {-1, -1, 1, 1};

static void physical_faces(const fclaw2d_hip_forest_t *f, int p, int phys[4])
{
    const int L = f->level[p];
    if (f->iso) {   /* every block is a domain of its own */
        const int w = 1 << L, lx = f->gx[p] & (w - 1), ly = f->gy[p] & (w - 1);
        phys[0] = !f->per_x && lx == 0;
        phys[1] = !f->per_x && lx == w - 1;
        phys[2] = !f->per_y && ly == 0;
        phys[3] = !f->per_y && ly == w - 1;
        return;
    }
    phys[0] = !f->per_x && f->gx[p] == 0;
    phys[1] = !f->per_x && f->gx[p] == (f->nbx << L) - 1;
    phys[2] = !f->per_y && f->gy[p] == 0;
    phys[3] = !f->per_y && f->gy[p] == (f->nby << L) - 1;
}

static void face_rel(const fclaw2d_hip_forest_t *f, int p, int iface, int mx, int my, nbinfo_t *r)
{
    const int L = f->level[p];
    const int64_t x = (int64_t) f->gx[p] + fdx[iface], y = (int64_t) f->gy[p] + fdy[iface];
    int64_t wx = x, wy = y;
    int q, k;
    r->rel = REL_BOUNDARY;
    if (!wrap(f, L, &wx, &wy, f->gx[p], f->gy[p])) return;
    if ((q = lookup(f, L, wx, wy)) >= 0) {
        r->rel = REL_SAME; r->nb[0] = q;
        r->ti[0] = -fdx[iface] * mx; r->tj[0] = -fdy[iface] * my;
        return;
    }
    if (L > 0 && (q = lookup(f, L - 1, wx >> 1, wy >> 1)) >= 0) {
        r->rel = REL_DOUBLE; r->nb[0] = q;
        return;
    }
    r->rel = REL_HALF;
    for (k = 0; k < 2; k++) {
        int64_t cx, cy;  /* unwrapped child coordinates adjacent to this face */
        if (iface == 0) { cx = 2 * x + 1; cy = 2 * (int64_t) f->gy[p] + k; }
        else if (iface == 1) { cx = 2 * x; cy = 2 * (int64_t) f->gy[p] + k; }
        else if (iface == 2) { cx = 2 * (int64_t) f->gx[p] + k; cy = 2 * y + 1; }
        else { cx = 2 * (int64_t) f->gx[p] + k; cy = 2 * y; }
        r->ti[k] = (int) ((2 * (int64_t) f->gx[p] - cx) * mx);
        r->tj[k] = (int) ((2 * (int64_t) f->gy[p] - cy) * my);
        if (!wrap(f, L + 1, &cx, &cy, 2 * (int64_t) f->gx[p], 2 * (int64_t) f->gy[p]) ||
            (q = lookup(f, L + 1, cx, cy)) < 0) { r->rel = REL_NONE; return; }
        r->nb[k] = q;
    }
}

static void corner_rel(const fclaw2d_hip_forest_t *f, int p, int ic, int mx, int my, nbinfo_t *r)
{
    const int L = f->level[p];
    const int64_t x = (int64_t) f->gx[p] + cdx[ic], y = (int64_t) f->gy[p] + cdy[ic];
    int64_t wx = x, wy = y;
    int q;
    r->rel = REL_NONE;
    if (!wrap(f, L, &wx, &wy, f->gx[p], f->gy[p])) return;
    if ((q = lookup(f, L, wx, wy)) >= 0) {
        r->rel = REL_SAME; r->nb[0] = q;
        r->ti[0] = -cdx[ic] * mx; r->tj[0] = -cdy[ic] * my;
        return;
    }
    if (L > 0 && (q = lookup(f, L - 1, wx >> 1, wy >> 1)) >= 0) {
        /* coarser diagonal leaf: a corner neighbour only if the corner point is
         * one of its corners; otherwise this corner is a hanging node */
        const int64_t px = (int64_t) f->gx[p] + (cdx[ic] > 0), py = (int64_t) f->gy[p] + (cdy[ic] > 0);
        const int64_t ox = 2 * (x >> 1), oy = 2 * (y >> 1);  /* unwrapped coarse origin, level L units */
        if ((px == ox || px == ox + 2) && (py == oy || py == oy + 2)) { r->rel = REL_DOUBLE; r->nb[0] = q; }
        return;
    }
    {
        int64_t cx = 2 * x + (cdx[ic] < 0), cy = 2 * y + (cdy[ic] < 0);
        r->ti[0] = (int) ((2 * (int64_t) f->gx[p] - cx) * mx);
        r->tj[0] = (int) ((2 * (int64_t) f->gy[p] - cy) * my);
        if (wrap(f, L + 1, &cx, &cy, 2 * (int64_t) f->gx[p], 2 * (int64_t) f->gy[p]) &&
            (q = lookup(f, L + 1, cx, cy)) >= 0) {
            r->rel = REL_HALF; r->nb[0] = q;
        }
    }
}

/* Relations of leaf p for patches of mx x my cells: [0..3] faces, [4..7]
 * corners; NULL if out of memory. */
static const nbinfo_t *relations(const fclaw2d_hip_forest_t *cf, int p, int mx, int my)
{
    fclaw2d_hip_forest_t *f = (fclaw2d_hip_forest_t *) cf;   /* the cache is not part of the forest's value */
    if (!f->rel_valid || f->rel_mx != mx || f->rel_my != my) {
        int q, k;
        if (f->rel_cap < (size_t) f->n || !f->rel) {
            free(f->rel);
            f->rel_cap = (size_t) f->n + (size_t) f->n / 4 + 16;
            f->rel = (nbinfo_t *) malloc(sizeof(nbinfo_t) * 8 * f->rel_cap);
            if (!f->rel) { f->rel_cap = 0; return NULL; }
        }
        for (q = 0; q < f->n; q++) {
            for (k = 0; k < 4; k++) face_rel(f, q, k, mx, my, &f->rel[8 * (size_t) q + k]);
            for (k = 0; k < 4; k++) corner_rel(f, q, k, mx, my, &f->rel[8 * (size_t) q + 4 + k]);
        }
        f->rel_mx = mx; f->rel_my = my; f->rel_valid = 1;
    }
    return f->rel + 8 * (size_t) p;
}

int fclaw2d_hip_forest_face_relation(const fclaw2d_hip_forest_t *f, int p, int iface, int nb[2])
{
    nbinfo_t r;
    if (!f || p < 0 || p >= f->n || iface < 0 || iface > 3) return -1;
    face_rel(f, p, iface, 1, 1, &r);
    if (nb) { nb[0] = (r.rel == REL_SAME || r.rel == REL_HALF || r.rel == REL_DOUBLE) ? r.nb[0] : -1;
              nb[1] = r.rel == REL_HALF ? r.nb[1] : -1; }
    return r.rel == REL_NONE ? -1 : r.rel;
}

/* ---------------------------------------------------------------- ghost program */

enum { OP_COPY_FACE, OP_COPY_CORNER, OP_AVG_FACE, OP_AVG_CORNER, OP_INTERP_FACE,
       OP_INTERP_CORNER, OP_PHYSBC };

typedef struct { int *ops; int cap; int n; } opbuf_t;

static void emit(opbuf_t *b, int type, int dst, int src, int which, int ti, int tj, int bc)
{
    if (b->n < b->cap) {
        int *o = b->ops + 8 * b->n;
        o[0] = type; o[1] = dst; o[2] = src; o[3] = which; o[4] = ti; o[5] = tj; o[6] = bc; o[7] = 0;
    }
    b->n++;
}

static int corner_interior(const int phys[4], int ic)
{
    const int fa = (ic & 1) ? 1 : 0, fb = (ic & 2) ? 3 : 2;
    return !phys[fa] && !phys[fb];
}

/* The leaves of each level in forest order: leaves byl[start[L] ..
 * start[L + 1]) have level L (levels 0..MAXL). */
enum { MAXL = 30 };
typedef struct { int start[MAXL + 2]; int *byl; } levels_t;

static int level_buckets(const fclaw2d_hip_forest_t *f, levels_t *lv)
{
    int p, L;
    lv->byl = (int *) malloc(sizeof(int) * ((size_t) f->n + 1));
    if (!lv->byl) return -1;
    memset(lv->start, 0, sizeof(lv->start));
    for (p = 0; p < f->n; p++) lv->start[f->level[p] + 1]++;
    for (L = 0; L <= MAXL; L++) lv->start[L + 1] += lv->start[L];
    {
        int fill[MAXL + 1];
        memcpy(fill, lv->start, sizeof(fill));
        for (p = 0; p < f->n; p++) lv->byl[fill[f->level[p]]++] = p;
    }
    return 0;
}

static void emit_physbc(const fclaw2d_hip_forest_t *f, const levels_t *lv, opbuf_t *b, int level,
                        const int mthbc[4], int tflag)
{
    int w, k;
    if (level < 0 || level > MAXL) return;
    for (w = lv->start[level]; w < lv->start[level + 1]; w++) {
        const int p = lv->byl[w];
        int phys[4], bc = 0, any = 0;
        physical_faces(f, p, phys);
        for (k = 0; k < 4; k++) {
            const int m = phys[k] ? mthbc[k] : -1;
            any |= phys[k];
            bc |= (m & 0xff) << (8 * k);
        }
        if (any) emit(b, OP_PHYSBC, p | tflag, 0, 0, 0, 0, bc);
    }
}

/* Face and corner exchanges of the patches at `level` (this side) with their
 * neighbours, as cb_face_fill / cb_corner_fill do for one exchange type.
 * `tflag` marks this side's slots as time-interpolated copies; for
 * interpolation the op writes the fine neighbour and reads this patch. */
static void emit_exchanges(const fclaw2d_hip_forest_t *f, const levels_t *lv, opbuf_t *b, int level,
                           int mx, int my, int kind, int tflag)
{
    int w, k;
    if (level < 0 || level > MAXL) return;
    for (w = lv->start[level]; w < lv->start[level + 1]; w++) {
        const int p = lv->byl[w];
        int phys[4];
        const nbinfo_t *rp;
        rp = relations(f, p, mx, my);
        physical_faces(f, p, phys);
        for (k = 0; k < 4; k++) {
            nbinfo_t r;
            if (phys[k]) continue;
            r = rp[k];
            if (kind == OP_COPY_FACE && r.rel == REL_SAME)
                emit(b, OP_COPY_FACE, p | tflag, r.nb[0] | tflag, k, r.ti[0], r.tj[0], 0);
            else if (kind == OP_AVG_FACE && r.rel == REL_HALF) {
                emit(b, OP_AVG_FACE, p | tflag, r.nb[0], k, r.ti[0], r.tj[0], 0);
                emit(b, OP_AVG_FACE, p | tflag, r.nb[1], k, r.ti[1], r.tj[1], 0);
            } else if (kind == OP_INTERP_FACE && r.rel == REL_HALF) {
                emit(b, OP_INTERP_FACE, r.nb[0], p | tflag, k, r.ti[0], r.tj[0], 0);
                emit(b, OP_INTERP_FACE, r.nb[1], p | tflag, k, r.ti[1], r.tj[1], 0);
            }
        }
        for (k = 0; k < 4; k++) {
            nbinfo_t r;
            if (!corner_interior(phys, k)) continue;
            r = rp[4 + k];
            if (kind == OP_COPY_FACE && r.rel == REL_SAME)
                emit(b, OP_COPY_CORNER, p | tflag, r.nb[0] | tflag, k, r.ti[0], r.tj[0], 0);
            else if (kind == OP_AVG_FACE && r.rel == REL_HALF)
                emit(b, OP_AVG_CORNER, p | tflag, r.nb[0], k, r.ti[0], r.tj[0], 0);
            else if (kind == OP_INTERP_FACE && r.rel == REL_HALF)
                emit(b, OP_INTERP_CORNER, r.nb[0], p | tflag, k, r.ti[0], r.tj[0], 0);
        }
    }
}

int fclaw2d_hip_forest_ghost_program_range(const fclaw2d_hip_forest_t *f, int mx, int my,
                                           int mbc, const int mthbc[4], int minlevel,
                                           int maxlevel, int time_interp, int *ops, int cap,
                                           int phase_start[7])
{
    opbuf_t b;
    levels_t lv;
    const int ti = minlevel - 1, TI = FCLAW2D_HIP_TIME_INTERP_SLOT;
    int L;
    (void) mbc;
    if (!f || f->n == 0 || !phase_start || !mthbc || minlevel > maxlevel) return -1;
    if (time_interp && ti < fclaw2d_hip_forest_min_level(f)) return -1;
    if (!relations(f, 0, mx, my)) return -1;   /* builds the relation cache used below */
    if (level_buckets(f, &lv)) return -1;
    b.ops = ops; b.cap = ops ? cap : 0; b.n = 0;

    /* phase 0: average fine -> coarse ghost cells (coarse levels, finest first) */
    phase_start[0] = b.n;
    for (L = maxlevel - 1; L >= minlevel; L--) emit_exchanges(f, &lv, &b, L, mx, my, OP_AVG_FACE, 0);
    if (time_interp) emit_exchanges(f, &lv, &b, ti, mx, my, OP_AVG_FACE, TI);
    /* phase 1: copy between same-size neighbours */
    phase_start[1] = b.n;
    for (L = maxlevel; L >= minlevel; L--) emit_exchanges(f, &lv, &b, L, mx, my, OP_COPY_FACE, 0);
    if (time_interp) emit_exchanges(f, &lv, &b, ti, mx, my, OP_COPY_FACE, TI);
    /* phase 2: physical BC on coarse levels */
    phase_start[2] = b.n;
    for (L = maxlevel - 1; L >= minlevel; L--) emit_physbc(f, &lv, &b, L, mthbc, 0);
    if (time_interp) emit_physbc(f, &lv, &b, ti, mthbc, TI);
    /* phase 3: interpolate coarse -> fine ghost cells */
    phase_start[3] = b.n;
    for (L = maxlevel - 1; L >= minlevel; L--) emit_exchanges(f, &lv, &b, L, mx, my, OP_INTERP_FACE, 0);
    if (time_interp) emit_exchanges(f, &lv, &b, ti, mx, my, OP_INTERP_FACE, TI);
    /* phase 4: physical BC on fine levels; phase 5: on all levels */
    phase_start[4] = b.n;
    for (L = maxlevel; L >= minlevel + 1; L--) emit_physbc(f, &lv, &b, L, mthbc, 0);
    phase_start[5] = b.n;
    for (L = maxlevel; L >= minlevel; L--) emit_physbc(f, &lv, &b, L, mthbc, 0);
    phase_start[6] = b.n;
    free(lv.byl);
    return b.n;
}

int fclaw2d_hip_forest_ghost_program(const fclaw2d_hip_forest_t *f, int mx, int my, int mbc,
                                     const int mthbc[4], int *ops, int cap, int phase_start[7])
{
    if (!f || f->n == 0) return -1;
    return fclaw2d_hip_forest_ghost_program_range(f, mx, my, mbc, mthbc,
                                                  fclaw2d_hip_forest_min_level(f),
                                                  fclaw2d_hip_forest_max_level(f), 0, ops, cap,
                                                  phase_start);
}

/* ---------------------------------------------------------------- sync program */

/* Time-synchronisation ops for fclaw2d_time_sync(minlevel, maxlevel) (see
 * hipclaw_set_sync_program_k), grouped by the patch they correct, in the
 * order correct_coarse_cells applies them (src/fclaw2d_time_sync.c:118-158
 * via cb_face_fill, src/fclaw2d_face_neighbors.c:321-383): same-size faces
 * 0..3, then faces with half-size neighbours, lower child first; patches
 * below minlevel are not corrected.  The resets (slot, face) are those of
 * fclaw2d_clawpatch_time_sync_reset (SAMESIZE at every level from minlevel,
 * F2C for coarse levels from minlevel and their fine neighbours, PHYS). */
int fclaw2d_hip_forest_sync_program_range(const fclaw2d_hip_forest_t *f, int mx, int my,
                                          int minlevel, int *ops, int cap, int *resets,
                                          int rcap, int *nresets)
{
    opbuf_t b;
    int p, k, g, nr = 0;
    if (!f || mx != my || cap < 0 || rcap < 0) return -1;
    if (f->n > 0 && !relations(f, 0, mx, my)) return -1;   /* builds the relation cache used below */
    b.ops = ops;
    b.cap = ops ? cap : 0;
    b.n = 0;
    for (p = 0; p < f->n; p++) {
        const nbinfo_t *rp;
        if (f->level[p] < minlevel) continue;
        rp = relations(f, p, mx, my);
        for (k = 0; k < 4; k++) {
            if (rp[k].rel == REL_NONE) return -1;
            if (rp[k].rel == REL_SAME) emit(&b, 0, p, rp[k].nb[0], k, 0, 0, 0);
        }
        for (k = 0; k < 4; k++) {
            if (rp[k].rel != REL_HALF) continue;
            for (g = 0; g < 2; g++) emit(&b, 1, p, rp[k].nb[g], k, rp[k].ti[g], rp[k].tj[g], 0);
        }
    }
    for (p = 0; p < f->n; p++) {
        const nbinfo_t *rp = relations(f, p, mx, my);
        for (k = 0; k < 4; k++) {
            int reset;
            const nbinfo_t r = rp[k];
            if (f->level[p] >= minlevel)
                reset = r.rel == REL_SAME || r.rel == REL_HALF || r.rel == REL_BOUNDARY ||
                        (r.rel == REL_DOUBLE && f->level[p] - 1 >= minlevel);
            else
                reset = 0;
            if (!reset) continue;
            if (resets && nr < rcap) { resets[2 * nr] = p; resets[2 * nr + 1] = k; }
            nr++;
        }
    }
    if (nresets) *nresets = nr;
    return b.n;
}

int fclaw2d_hip_forest_sync_program(const fclaw2d_hip_forest_t *f, int mx, int my, int *ops,
                                    int cap)
{
    if (!f || f->n == 0) return -1;
    return fclaw2d_hip_forest_sync_program_range(f, mx, my, fclaw2d_hip_forest_min_level(f),
                                                 ops, cap, NULL, 0, NULL);
}
