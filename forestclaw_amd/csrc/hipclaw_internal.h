// Internal interfaces between the C-ABI layer (hipclaw_capi.hip) and the two
// kernel translation units (FP64 parity form, FP32-bulk form).
#pragma once

#include <atomic>

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "../../include/hipclaw.h"

#include "hipclaw_launch.h"

namespace hipclaw {

// One ghost-fill operation; every op touches only ghost cells of `dst`.
// Layout is shared with the host op-list builder (int32 x 8).
struct GhostOp {
    int type;   // GHOST_* below
    int dst;    // slot written
    int src;    // slot read (unused for physical BC)
    int which;  // face or corner of `dst` (copy) / of the coarse patch (average, interpolate)
    int ti, tj; // index translation (see the op builder)
    int bc;     // physical BC: mthbc of faces 0..3 packed as 4 signed bytes
    int reps;   // physical BC: times the BC is applied in a row (the library sets it)
};

enum {
    GHOST_COPY_FACE = 0,
    GHOST_COPY_CORNER = 1,
    GHOST_AVERAGE_FACE = 2,
    GHOST_AVERAGE_CORNER = 3,
    GHOST_INTERP_FACE = 4,
    GHOST_INTERP_CORNER = 5,
    GHOST_PHYSBC = 6,
};

struct GhostArgs {
    const GhostOp* ops;
    int nops;
    double* q;
    double* qti;           // time-interpolated copies (slots flagged HIPCLAW_TIME_INTERP_SLOT)
    long long pstride;
    int mx, my, mbc, meqn;
};

// One time-synchronisation op (int32 x 8), grouped by destination slot.
enum { SYNC_SAMESIZE = 0, SYNC_F2C = 1,
       // fused step, FP32 flux form only: a same-size op that also writes its
       // partner's increment (dst's face is src's face ^ 1; both are the same
       // value there, since both read the same two shared-flux registers)
       SYNC_SAMESIZE_PAIR = 2 };

struct SyncArgs {
    const int* ops;      // [nops][8]: type, dst, src, face, ti, tj, -, -
    const int* start;    // [ngroups + 1]: ops of group g are start[g] .. start[g+1]-1
    int ngroups;
    int maxgroup;        // most ops in one group
    double* q;
    const double* reg;   // flux registers, layout of StepArgs::reg
    const double* dxs;
    const double* dys;
    long long pstride, rstride;
    const int* resets;   // [nresets][2]: (slot, face) registers zeroed after the corrections
    int nresets;
    int mx, mbc, meqn;   // mx == my
    int fluxreg;         // registers hold one shared edge flux (FP32 flux form)
};

// Fused step (ghost layout 2).  What fills the ghost strip of each face of a
// patch, taken from its ghost ops: every ghost cell the fused launch needs is
// recomputed from interior cells through this table, never read from a ghost
// cell another wave writes.
enum { FREL_NONE = 0, FREL_SAME = 1, FREL_HALF = 2, FREL_DOUBLE = 3, FREL_PHYS = 4 };
struct FaceRel {
    int rel;
    int nb0, ti0, tj0;   // same size: the neighbour; half size: the first child;
                         // double size: the coarse patch and its interpolation's (ti, tj)
    int nb1, ti1, tj1;   // half size: the second child
    int bc;              // physical: mthbc of this face (1 or 3); double size: the
                         // coarse patch's face toward this one
};

// One launch of the fused ghost fill and time-sync increments: ghost ops
// [ncorner0, ncorner0 + nbig) of g.ops one wave each, the corner ops
// [0, ncorner0) and [ncorner0 + nbig, ... + ncorner1) four to a wave, then
// sync ops [0, nsync) of sops one wave each.
struct FlatArgs {
    GhostArgs g;
    int ncorner0, nbig, ncorner1;
    const FaceRel* frel;       // [slot][4]
    const int* sops;           // [nsync][8] as SyncArgs::ops
    int nsync;
    const double* reg;
    const double* dxs;
    const double* dys;
    long long rstride;
    double* qinc;              // [slot][face][m][pos] time-sync increments
    long long cstride;
    int fluxreg;
};

// Regrid op kinds (the forest's provenance kinds).
enum { REGRID_SAME = 0, REGRID_CHILD = 1, REGRID_PARENT = 2 };

// Launchers (return hipError_t as int).
int launch_tag_patches(const double* q, const int* slots, int n, const double* dxs,
                       const double* dys, long long pstride, int mx, int my, int mbc, int ivar,
                       double thr_refine, double thr_coarsen, int* out, hipStream_t st);
int launch_regrid_remap(const double* qold, double* qnew, const int* ops, int nops,
                        long long pstride, int mx, int my, int mbc, int meqn, hipStream_t st);
int launch_time_sync(const SyncArgs& a, hipStream_t st);
// Exchange ops (copy, average, interpolate) of one phase, one wave per op.
int launch_ghost_exchange(const GhostArgs& g, hipStream_t st);
// Physical-BC ops, one workgroup per patch.
int launch_ghost_physbc(const GhostArgs& g, hipStream_t st);
// Per-destination groups: a patch's exchange ops, then its physical BC;
// groups[b] = (first op, end, physical-BC op or -1, -).
int launch_ghost_groups(const GhostArgs& g, const int4* groups, int ngroups, hipStream_t st);
// The fused ghost fill (+ time-sync increments into qinc) of one step.
int launch_ghost_flat(const FlatArgs& f, hipStream_t st);
// qcorr -> the interior edge cells of q (slots 0..n-1, faces in cmask).
int launch_apply_qcorr(double* q, const double* qcorr, const int* cmask, const int* slots, int n,
                       long long pstride, long long cstride, int mx, int my, int mbc, int meqn,
                       hipStream_t st);
int launch_step_fp64(int solver, int tile, int waves, const StepArgs& a, const StepLaunch& st);
int launch_step_fp32(int solver, int tile, int waves, const StepArgs& a, const StepLaunch& st);
// The default waves per workgroup of a tile (0 where there is no choice).
int default_waves_fp64(int solver, int tile);
int default_waves_fp32(int solver, int tile);
int launch_reset_scalar(unsigned long long* p, hipStream_t st);
// Whole patch slabs src[slot] -> dst[slot].
int launch_copy_patches(const double* src, double* dst, const int* slots, int n, long long pstride,
                        hipStream_t st);
// qti = qlast + alpha (q - qlast) on the interior band of width mbc.
int launch_ghost_message(double* q, double* reg, const int* slots, int n, const int* band,
                         int nband, long long pstride, long long rstride, int mx, int my, int mbc,
                         int meqn, const double* dxs, const double* dys, double* buf,
                         long long elems, int unpack, hipStream_t st);
int launch_vtk_cell_values(const double* q, int first, int count, long long pstride, int mx, int my,
                           int mbc, int meqn, float* out, hipStream_t st);
int launch_timeinterp(const double* q, const double* qlast, double* qti, const int* slots, int n,
                      long long pstride, int mx, int my, int mbc, int meqn, double alpha,
                      hipStream_t st);
int launch_copy_interior(const double* src, double* dst, const int* slots, int n,
                         long long pstride, int mx, int my, int mbc, int meqn,
                         hipStream_t st);
// Per-slot compensated FP64 sums of q*dx*dy over interior cells: out[slot][m] = (hi, lo)
int launch_mass(const double* q, const int* slots, int n, const double* dxs, const double* dys,
                long long pstride, int mx, int my, int mbc, int meqn, double* out,
                hipStream_t st);

// Largest tile the FP64 / FP32 kernels are instantiated for, per solver.
bool step_supported(int solver, int precision, int tile);

}  // namespace hipclaw
