/*
 * Minimal host forest for the batched patch-update path: the leaves of a
 * forest of quadtrees over a brick of nbx x nby unit blocks (optionally
 * periodic), kept 2:1 balanced across faces and corners, in p4est's order
 * (block, then Morton index).  It stands in for the topology queries the
 * reference takes from p4est (src/forestclaw2d.c: patch face/corner neighbours
 * and transforms) and produces the device ghost-fill program.
 *
 * A patch is identified by (level, gx, gy): its column and row at its level
 * counted over the whole brick (block bx holds columns bx*2^level ..).
 */
#ifndef FCLAW2D_HIP_FOREST_H
#define FCLAW2D_HIP_FOREST_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fclaw2d_hip_forest fclaw2d_hip_forest_t;

/* Uniform forest at `minlevel` over nbx x nby blocks. */
fclaw2d_hip_forest_t *fclaw2d_hip_forest_new(int nbx, int nby, int periodic_x,
                                             int periodic_y, int minlevel);
/* Forest from an explicit leaf list (must tile the brick); it is balanced if
 * `balance` is nonzero.  Returns NULL on an invalid leaf set. */
fclaw2d_hip_forest_t *fclaw2d_hip_forest_from_leaves(int nbx, int nby, int periodic_x,
                                                     int periodic_y, int n, const int *level,
                                                     const int *gx, const int *gy,
                                                     int balance);
void fclaw2d_hip_forest_destroy(fclaw2d_hip_forest_t *f);

/* Make every block of the brick a domain of its own (an ensemble of runs in
 * one forest): no neighbour relations, ghost exchange or 2:1 balance cross a
 * block boundary, and block faces are physical boundaries (or wrap around
 * their own block when periodic).  Set it before refining.  Returns 0. */
int fclaw2d_hip_forest_set_isolated(fclaw2d_hip_forest_t *f, int isolated);

int fclaw2d_hip_forest_num_patches(const fclaw2d_hip_forest_t *f);
int fclaw2d_hip_forest_min_level(const fclaw2d_hip_forest_t *f);
int fclaw2d_hip_forest_max_level(const fclaw2d_hip_forest_t *f);
/* Copy out the leaves in forest order. */
void fclaw2d_hip_forest_leaves(const fclaw2d_hip_forest_t *f, int *level, int *gx, int *gy);

/* Refine every flagged leaf (flags[p] != 0) below maxlevel by one level, then
 * restore the 2:1 face and corner balance by refining further leaves.
 * Returns the new number of leaves, or -1 on error. */
int fclaw2d_hip_forest_refine(fclaw2d_hip_forest_t *f, const int *flags, int maxlevel);

/* One regrid step of the forest (the adapt of fclaw2d_regrid,
 * src/fclaw2d_regrid.c:235-330): leaves flagged in refine[] (below maxlevel)
 * are split, the forest is rebalanced, then every family of four siblings all
 * flagged in coarsen[] (none refined, above minlevel) is replaced by its
 * parent where that keeps the 2:1 balance.  Returns the new number of leaves
 * or -1.  The provenance of each new leaf is then available (below). */
int fclaw2d_hip_forest_adapt(fclaw2d_hip_forest_t *f, const int *refine, const int *coarsen,
                             int minlevel, int maxlevel);

enum {
    FCLAW2D_HIP_REGRID_SAME = 0,    /* src[0] = old leaf index */
    FCLAW2D_HIP_REGRID_CHILD = 1,   /* src[0] = old parent, src[1] = child index 0..3 */
    FCLAW2D_HIP_REGRID_PARENT = 2   /* src[0..3] = old children in z-order */
};
/* Provenance of the leaves after the last fclaw2d_hip_forest_adapt:
 * kind[n] and src[4 n]; returns n or -1. */
int fclaw2d_hip_forest_provenance(const fclaw2d_hip_forest_t *f, int *kind, int *src);

/* Ghost-fill program for the current leaves (see hipclaw_set_ghost_program).
 * mthbc gives the physical BC of the four domain faces.  Writes up to `cap`
 * ops of 8 ints and phase_start[7]; returns the number of ops needed (call
 * with cap = 0 to size the buffer), or -1 on error. */
int fclaw2d_hip_forest_ghost_program(const fclaw2d_hip_forest_t *f, int mx, int my, int mbc,
                                     const int mthbc[4], int *ops, int cap,
                                     int phase_start[7]);

/* Slot flag in ghost-fill ops: the slot's time-interpolated copy (see
 * hipclaw_timeinterp) rather than its solution. */
#define FCLAW2D_HIP_TIME_INTERP_SLOT (1 << 30)

/* Ghost-fill program of fclaw2d_ghost_update(minlevel, maxlevel, time_interp)
 * (src/fclaw2d_ghost_fill.c:611-830): with time_interp, level minlevel-1
 * takes part through its time-interpolated copies (ops flagged with
 * FCLAW2D_HIP_TIME_INTERP_SLOT).  Same conventions as
 * fclaw2d_hip_forest_ghost_program, which is the full range without it. */
int fclaw2d_hip_forest_ghost_program_range(const fclaw2d_hip_forest_t *f, int mx, int my,
                                           int mbc, const int mthbc[4], int minlevel,
                                           int maxlevel, int time_interp, int *ops, int cap,
                                           int phase_start[7]);

/* Time-synchronisation program of fclaw2d_time_sync(minlevel, maxlevel):
 * returns the number of ops (up to `cap` written) and the (slot, face)
 * register resets in `resets` (up to rcap pairs; *nresets = number needed). */
int fclaw2d_hip_forest_sync_program_range(const fclaw2d_hip_forest_t *f, int mx, int my,
                                          int minlevel, int *ops, int cap, int *resets,
                                          int rcap, int *nresets);

/* Time-synchronisation program for the current leaves (see
 * hipclaw_set_sync_program): writes up to `cap` ops of 8 ints and returns the
 * number needed (cap = 0 to size the buffer), or -1 on error (mx != my). */
int fclaw2d_hip_forest_sync_program(const fclaw2d_hip_forest_t *f, int mx, int my, int *ops,
                                    int cap);

/* Face neighbour relation of patch p (0 boundary, 1 same size, 2 half size,
 * 3 double size) -- exposed for tests and the host plugin layer. */
int fclaw2d_hip_forest_face_relation(const fclaw2d_hip_forest_t *f, int p, int iface,
                                     int nb[2]);

#ifdef __cplusplus
}
#endif
#endif
