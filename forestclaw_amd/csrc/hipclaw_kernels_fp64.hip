// FP64 kernels: the oracle-parity patch step, the ghost fill, and the
// compensated mass sums.  This translation unit is compiled with
// -ffp-contract=off so that every FP64 operation rounds exactly as the
// reference's CPU Fortran does.
#include "hipclaw_internal.h"

#include <algorithm>

namespace hipclaw {

#define QIX(i, j, m) (((long long) (m) * nyp + ((j) + mbc - 1)) * nxp + ((i) + mbc - 1))

// fclaw2d_clawpatch_compute_slopes with mth = 5
// (src/patches/clawpatch/fclaw2d_clawpatch_utils.f:45-55)
__device__ __forceinline__ double amr_slope(double sl, double sr) {
    const double sc = (sl + sr) / 2.0;
    const double mn = fmin(fmin(2 * fabs(sl), 2 * fabs(sr)), fabs(sc));
    return mn * fmax(0.0, copysign(1.0, sl * sr)) * copysign(1.0, sc);
}

__device__ __forceinline__ void corner_cell(int mx, int my, int icorner, int ibc, int jbc,
                                            int& i1, int& j1) {
    i1 = (icorner & 1) ? mx + ibc : 1 - ibc;
    j1 = (icorner & 2) ? my + jbc : 1 - jbc;
}

// Interpolate one coarse cell (ic,jc) of `qc` into the 2x2 fine cells whose first
// child is (fi,fj) in `qf` (fclaw2d_clawpatch46_interpolate.f:120-137).
__device__ __forceinline__ void interp_cell(const double* qc, double* qf, int m, int ic, int jc,
                                            int fi, int fj, int nxp, int nyp, int mbc) {
    const double c = qc[QIX(ic, jc, m)];
    const double gx = amr_slope(c - qc[QIX(ic - 1, jc, m)], qc[QIX(ic + 1, jc, m)] - c);
    const double gy = amr_slope(c - qc[QIX(ic, jc - 1, m)], qc[QIX(ic, jc + 1, m)] - c);
    qf[QIX(fi, fj, m)] = c + gx * -0.25 + gy * -0.25;
    qf[QIX(fi + 1, fj, m)] = c + gx * 0.25 + gy * -0.25;
    qf[QIX(fi, fj + 1, m)] = c + gx * -0.25 + gy * 0.25;
    qf[QIX(fi + 1, fj + 1, m)] = c + gx * 0.25 + gy * 0.25;
}

// One wave per ghost op, OPW ops (waves) per workgroup.  Ops of one phase
// write disjoint ghost cells and read only cells no op of the phase writes,
// so each lane issues the loads of up to U elements before their stores: the
// source and destination slabs are the same allocation, and without this the
// compiler must wait for every store before the next load.
// W lanes (t = 0..W-1) work on the op: a whole wave, or a quarter of one for
// the corner ops of the fused launch.
template <int W = 64>
__device__ __forceinline__ void ghost_exchange_op(const GhostArgs& g, const GhostOp& op, int t) {
    const int mx = g.mx, my = g.my, mbc = g.mbc, meqn = g.meqn;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    const int TI = HIPCLAW_TIME_INTERP_SLOT;
    double* qd = ((op.dst & TI) ? g.qti : g.q) + (long long) (op.dst & ~TI) * g.pstride;
    const double* qs = ((op.src & TI) ? g.qti : g.q) + (long long) (op.src & ~TI) * g.pstride;
    constexpr int U = 4;

    switch (op.type) {
    case GHOST_COPY_FACE:      // fclaw2d_clawpatch46_copy.f:40-84
    case GHOST_COPY_CORNER: {  // fclaw2d_clawpatch46_copy.f:95-133
        const bool face = op.type == GHOST_COPY_FACE;
        const int len = (op.which < 2) ? my : mx;
        const int n = face ? meqn * mbc * len : meqn * mbc * mbc;
        for (int base = t; base < n; base += W * U) {
            double v[U];
            long long di[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = base + W * u;
                di[u] = -1;
                if (e >= n) continue;
                int m, i1, j1;
                if (face) {
                    m = e / (mbc * len);
                    const int r = e - m * mbc * len, l = r / mbc, bb = r - l * mbc;  // ghost layer 0..mbc-1
                    if (op.which == 0) { i1 = 1 - mbc + bb; j1 = 1 + l; }
                    else if (op.which == 1) { i1 = mx + 1 + bb; j1 = 1 + l; }
                    else if (op.which == 2) { i1 = 1 + l; j1 = 1 - mbc + bb; }
                    else { i1 = 1 + l; j1 = my + 1 + bb; }
                } else {
                    m = e / (mbc * mbc);
                    const int r = e - m * mbc * mbc;
                    corner_cell(mx, my, op.which, r / mbc + 1, r % mbc + 1, i1, j1);
                }
                v[u] = qs[QIX(i1 + op.ti, j1 + op.tj, m)];
                di[u] = QIX(i1, j1, m);
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (di[u] >= 0) qd[di[u]] = v[u];
        }
        break;
    }
    case GHOST_AVERAGE_FACE:      // fclaw2d_clawpatch46_average.f:61-174 (manifold = 0)
    case GHOST_AVERAGE_CORNER: {  // fclaw2d_clawpatch46_average.f:225-270
        const bool face = op.type == GHOST_AVERAGE_FACE;
        const int len = (op.which < 2) ? my : mx;
        const int n = face ? meqn * mbc * len : meqn * mbc * mbc;
        for (int base = t; base < n; base += W * U) {
            double v[U];
            long long di[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = base + W * u;
                di[u] = -1;
                if (e >= n) continue;
                int m, ic, jc;
                if (face) {
                    m = e / (mbc * len);
                    const int r = e - m * mbc * len, l = r / mbc, bb = r - l * mbc + 1;  // ibc/jbc 1..mbc
                    if (op.which < 2) { ic = (op.which == 0) ? 1 - bb : mx + bb; jc = 1 + l; }
                    else { ic = 1 + l; jc = (op.which == 2) ? 1 - bb : my + bb; }
                } else {
                    m = e / (mbc * mbc);
                    const int r = e - m * mbc * mbc;
                    corner_cell(mx, my, op.which, r / mbc + 1, r % mbc + 1, ic, jc);
                }
                const int fi = 2 * ic - 1 + op.ti, fj = 2 * jc - 1 + op.tj;
                // faces: only cells whose four children are interior to this fine grid
                if (face && (fi < 1 || fi + 1 > mx || fj < 1 || fj + 1 > my)) continue;
                double sum = 0;
                sum = sum + qs[QIX(fi, fj, m)];
                sum = sum + qs[QIX(fi + 1, fj, m)];
                sum = sum + qs[QIX(fi, fj + 1, m)];
                sum = sum + qs[QIX(fi + 1, fj + 1, m)];
                v[u] = sum / 4.0;
                di[u] = QIX(ic, jc, m);
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (di[u] >= 0) qd[di[u]] = v[u];
        }
        break;
    }
    case GHOST_INTERP_FACE:      // fclaw2d_clawpatch46_interpolate.f:91-197; dst = fine, src = coarse
    case GHOST_INTERP_CORNER: {  // fclaw2d_clawpatch46_interpolate.f:263-311
        const bool face = op.type == GHOST_INTERP_FACE;
        const int hb = mbc / 2;
        const int n = face ? meqn * hb * mx : meqn * hb * hb;  // the reference bounds the face loop by mx
        constexpr int UI = 2;
        for (int base = t; base < n; base += W * UI) {
            double c[UI], gx[UI], gy[UI];
            long long d0[UI];
#pragma unroll
            for (int u = 0; u < UI; u++) {
                const int e = base + W * u;
                d0[u] = -1;
                if (e >= n) continue;
                int m, ic, jc;
                if (face) {
                    m = e / (hb * mx);
                    const int r = e - m * hb * mx, l = r / hb, bb = r - l * hb + 1;
                    if (op.which < 2) { ic = (op.which == 0) ? bb : mx - bb + 1; jc = 1 + l; }
                    else { ic = 1 + l; jc = (op.which == 2) ? bb : my - bb + 1; }
                } else {
                    m = e / (hb * hb);
                    const int r = e - m * hb * hb, ibc = r / hb + 1, jbc = r % hb + 1;
                    ic = (op.which & 1) ? mx - ibc + 1 : ibc;
                    jc = (op.which & 2) ? my - jbc + 1 : jbc;
                }
                const int fi = 2 * ic - 1 + op.ti, fj = 2 * jc - 1 + op.tj;
                if (face && (fi < 1 - mbc || fi + 1 > mx + mbc || fj < 1 - mbc || fj + 1 > my + mbc)) continue;
                const double cc = qs[QIX(ic, jc, m)];
                gx[u] = amr_slope(cc - qs[QIX(ic - 1, jc, m)], qs[QIX(ic + 1, jc, m)] - cc);
                gy[u] = amr_slope(cc - qs[QIX(ic, jc - 1, m)], qs[QIX(ic, jc + 1, m)] - cc);
                c[u] = cc;
                d0[u] = QIX(fi, fj, m);
            }
            // fine cells (fi..fi+1, fj..fj+1), fclaw2d_clawpatch46_interpolate.f:120-137
#pragma unroll
            for (int u = 0; u < UI; u++) {
                if (d0[u] < 0) continue;
                qd[d0[u]] = c[u] + gx[u] * -0.25 + gy[u] * -0.25;
                qd[d0[u] + 1] = c[u] + gx[u] * 0.25 + gy[u] * -0.25;
                qd[d0[u] + nxp] = c[u] + gx[u] * -0.25 + gy[u] * 0.25;
                qd[d0[u] + nxp + 1] = c[u] + gx[u] * 0.25 + gy[u] * 0.25;
            }
        }
        break;
    }
    default:
        break;
    }
}

template <int OPW>
__global__ void __launch_bounds__(64 * OPW) ghost_phase(GhostArgs g) {
    const int o = blockIdx.x * OPW + (OPW > 1 ? (int) (threadIdx.x >> 6) : 0);
    if (o >= g.nops) return;
    ghost_exchange_op(g, g.ops[o], threadIdx.x & 63);
}

// Physical BC of one patch per workgroup (clawpack46_bc2_default.f): faces 0
// and 1 together, then faces 2 and 3, whose strips include the corner cells
// the first two wrote.  Within a pair no element reads a cell the pair
// writes (mx, my >= 2 mbc), so all loads of a lane precede its stores.  The
// BC reads and writes cells of its own patch only, so applying it op.reps
// times in a row equals that many phases of it.
constexpr int PHYSBC_THREADS = 256;
// Every thread of the workgroup (PHYSBC_THREADS) calls this: it synchronises.
__device__ __forceinline__ void physbc_op(const GhostArgs& g, const GhostOp& op) {
    const int mx = g.mx, my = g.my, mbc = g.mbc, meqn = g.meqn;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    const int TI = HIPCLAW_TIME_INTERP_SLOT;
    double* qd = ((op.dst & TI) ? g.qti : g.q) + (long long) (op.dst & ~TI) * g.pstride;
    constexpr int U = 4;
    const int npass = 2 * (op.reps > 1 ? op.reps : 1);
    for (int pass = 0; pass < npass; pass++) {
        const int pair = pass & 1;
        const int len = pair == 0 ? my + 2 * mbc : mx + 2 * mbc;
        const int per = meqn * mbc * len, n = 2 * per;
        for (int base = threadIdx.x; base < n; base += PHYSBC_THREADS * U) {
            double v[U];
            long long di[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int e = base + PHYSBC_THREADS * u;
                di[u] = -1;
                if (e >= n) continue;
                const int face = 2 * pair + (e >= per);
                const int mth = (signed char) ((op.bc >> (8 * face)) & 0xff);
                if (mth != 1 && mth != 3) continue;
                const int f = e - (e >= per ? per : 0);
                const int m = f / (mbc * len), r = f - m * mbc * len;
                const int l = r / mbc, b = r - l * mbc + 1;
                int id, jd, is, js;
                if (face < 2) {
                    jd = js = 1 - mbc + l;
                    if (face == 0) { id = 1 - b; is = (mth == 1) ? 1 : b; }
                    else { id = mx + b; is = (mth == 1) ? mx : mx + 1 - b; }
                } else {
                    id = is = 1 - mbc + l;
                    if (face == 2) { jd = 1 - b; js = (mth == 1) ? 1 : b; }
                    else { jd = my + b; js = (mth == 1) ? my : my + 1 - b; }
                }
                const int mn = (face < 2) ? 1 : 2;  // normal momentum negated on walls
                const double x = qd[QIX(is, js, m)];
                v[u] = (mth == 3 && m == mn) ? -x : x;
                di[u] = QIX(id, jd, m);
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (di[u] >= 0) qd[di[u]] = v[u];
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(PHYSBC_THREADS) physbc_phase(GhostArgs g) {
    const int o = blockIdx.x;
    if (o >= g.nops) return;
    physbc_op(g, g.ops[o]);
}

// One workgroup per destination patch of a phase group: the patch's
// exchange ops (one wave each, the four waves taking them in turn), then,
// once they are all written, its physical BC.  groups[b] = (first exchange
// op, end of its exchange ops, physical-BC op or -1, unused), indices into
// g.ops.  A group
// writes ghost cells of its own patch only; exchange ops read interior cells
// of other patches, or (interpolation) cells of a coarser patch that its own
// physical BC may write again with the same values (the BC is a function of
// the patch's interior and its non-physical ghost cells).
__global__ void __launch_bounds__(PHYSBC_THREADS) ghost_groups(GhostArgs g, const int4* __restrict__ groups,
                                                              int ngroups) {
    const int b = blockIdx.x;
    if (b >= ngroups) return;
    const int4 gr = groups[b];
    const int w = threadIdx.x >> 6;
    for (int o = gr.x + w; o < gr.y; o += PHYSBC_THREADS / 64) ghost_exchange_op(g, g.ops[o], threadIdx.x & 63);
    if (gr.z >= 0) {
        __syncthreads();
        physbc_op(g, g.ops[gr.z]);
    }
}

// ---------------------------------------------------------------- fused ghost fill + sync
// Cell (i, j) of patch `slot` as the ghost fill of this step leaves it, for an
// interior cell or a face ghost cell that is not a corner and not filled by
// interpolation: the neighbour's interior cell (same size), the average of the
// fine child's four cells (fclaw2d_clawpatch46_average.f:61-174, the child
// whose cells lie in its interior), or the physical BC of the patch's own
// interior (clawpack46_bc2_default.f).  All of it reads interior cells only,
// which no wave of the fused launch writes.
__device__ double fused_cell_simple(const GhostArgs& g, const FaceRel* __restrict__ frel, int slot, int i,
                                    int j, int m) {
    const int mx = g.mx, my = g.my, mbc = g.mbc;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    const double* qc = g.q + (long long) slot * g.pstride;
    const int k = i < 1 ? 0 : i > mx ? 1 : j < 1 ? 2 : j > my ? 3 : -1;
    if (k < 0) return qc[QIX(i, j, m)];
    const FaceRel r = frel[4 * slot + k];
    if (r.rel == FREL_SAME) return g.q[(long long) r.nb0 * g.pstride + QIX(i + r.ti0, j + r.tj0, m)];
    if (r.rel == FREL_HALF) {
        int fi = 2 * i - 1 + r.ti0, fj = 2 * j - 1 + r.tj0, nb = r.nb0;
        if (fi < 1 || fi + 1 > mx || fj < 1 || fj + 1 > my) {
            fi = 2 * i - 1 + r.ti1;
            fj = 2 * j - 1 + r.tj1;
            nb = r.nb1;
            // one of the two children covers every cell of the face strip
            if (fi < 1 || fi + 1 > mx || fj < 1 || fj + 1 > my) return 0.0;
        }
        const double* qs = g.q + (long long) nb * g.pstride;
        double sum = 0;
        sum = sum + qs[QIX(fi, fj, m)];
        sum = sum + qs[QIX(fi + 1, fj, m)];
        sum = sum + qs[QIX(fi, fj + 1, m)];
        sum = sum + qs[QIX(fi + 1, fj + 1, m)];
        return sum / 4.0;
    }
    // FREL_PHYS (the host admits nothing else where this is called): extrapolation
    // copies the first interior cell, a wall mirrors the strip and negates the
    // normal momentum
    const int mth = r.bc;
    int is = i, js = j, mn;
    if (k == 0) { is = mth == 1 ? 1 : 1 - i; mn = 1; }
    else if (k == 1) { is = mth == 1 ? mx : 2 * mx + 1 - i; mn = 1; }
    else if (k == 2) { js = mth == 1 ? 1 : 1 - j; mn = 2; }
    else { js = mth == 1 ? my : 2 * my + 1 - j; mn = 2; }
    const double x = qc[QIX(is, js, m)];
    return (mth == 3 && m == mn) ? -x : x;
}

// Interpolated value of coarse cell (ic, jc) of patch c at child (a, b) of its
// 2 x 2 fine cells (fclaw2d_clawpatch46_interpolate.f:120-137), the coarse
// neighbours read through fused_cell_simple.
__device__ __forceinline__ double interp_child(const GhostArgs& g, const FaceRel* __restrict__ frel, int c,
                                               int ic, int jc, int m, int a, int b) {
    const double cc = fused_cell_simple(g, frel, c, ic, jc, m);
    const double gx = amr_slope(cc - fused_cell_simple(g, frel, c, ic - 1, jc, m),
                                fused_cell_simple(g, frel, c, ic + 1, jc, m) - cc);
    const double gy = amr_slope(cc - fused_cell_simple(g, frel, c, ic, jc - 1, m),
                                fused_cell_simple(g, frel, c, ic, jc + 1, m) - cc);
    return cc + gx * (a ? 0.25 : -0.25) + gy * (b ? 0.25 : -0.25);
}

// fused_cell_simple, plus the face strips filled by interpolation from a coarser
// patch (recomputed from the coarse patch's cells).
__device__ double fused_cell(const GhostArgs& g, const FaceRel* __restrict__ frel, int slot, int i, int j,
                             int m) {
    const int mx = g.mx, my = g.my;
    const int k = i < 1 ? 0 : i > mx ? 1 : j < 1 ? 2 : j > my ? 3 : -1;
    if (k >= 0) {
        const FaceRel r = frel[4 * slot + k];
        if (r.rel == FREL_DOUBLE) {
            // fine cell (i, j) = child (a, b) of coarse cell (ic, jc):
            // i = 2 ic - 1 + ti + a
            const int u = i - r.ti0 + 1, v = j - r.tj0 + 1;
            return interp_child(g, frel, r.nb0, u >> 1, v >> 1, m, u & 1, v & 1);
        }
    }
    return fused_cell_simple(g, frel, slot, i, j, m);
}

// Interpolation op (dst = fine, src = coarse) as in ghost_exchange_op, its
// coarse-side reads through fused_cell_simple.
template <int W = 64>
__device__ void fused_interp_op(const GhostArgs& g, const FaceRel* __restrict__ frel, const GhostOp& op,
                                int t) {
    const int mx = g.mx, my = g.my, mbc = g.mbc, meqn = g.meqn;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    double* qd = g.q + (long long) op.dst * g.pstride;
    const bool face = op.type == GHOST_INTERP_FACE;
    const int hb = mbc / 2;
    const int n = face ? meqn * hb * mx : meqn * hb * hb;
    for (int e = t; e < n; e += W) {
        int m, ic, jc;
        if (face) {
            m = e / (hb * mx);
            const int r = e - m * hb * mx, l = r / hb, bb = r - l * hb + 1;
            if (op.which < 2) { ic = (op.which == 0) ? bb : mx - bb + 1; jc = 1 + l; }
            else { ic = 1 + l; jc = (op.which == 2) ? bb : my - bb + 1; }
        } else {
            m = e / (hb * hb);
            const int r = e - m * hb * hb, ibc = r / hb + 1, jbc = r % hb + 1;
            ic = (op.which & 1) ? mx - ibc + 1 : ibc;
            jc = (op.which & 2) ? my - jbc + 1 : jbc;
        }
        const int fi = 2 * ic - 1 + op.ti, fj = 2 * jc - 1 + op.tj;
        if (face && (fi < 1 - mbc || fi + 1 > mx + mbc || fj < 1 - mbc || fj + 1 > my + mbc)) continue;
        const int c = op.src;
        const double cc = fused_cell_simple(g, frel, c, ic, jc, m);
        const double gx = amr_slope(cc - fused_cell_simple(g, frel, c, ic - 1, jc, m),
                                    fused_cell_simple(g, frel, c, ic + 1, jc, m) - cc);
        const double gy = amr_slope(cc - fused_cell_simple(g, frel, c, ic, jc - 1, m),
                                    fused_cell_simple(g, frel, c, ic, jc + 1, m) - cc);
        const long long d0 = QIX(fi, fj, m);
        qd[d0] = cc + gx * -0.25 + gy * -0.25;
        qd[d0 + 1] = cc + gx * 0.25 + gy * -0.25;
        qd[d0 + nxp] = cc + gx * -0.25 + gy * 0.25;
        qd[d0 + nxp + 1] = cc + gx * 0.25 + gy * 0.25;
    }
}

// Physical BC of one patch by one wave (clawpack46_bc2_default.f), each ghost
// cell written once with its final value: the sources are interior cells or
// ghost cells of a neighbouring face strip, recomputed through fused_cell, and
// a corner both of whose faces are physical takes the y-face's value (the
// reference applies faces 2 and 3 after faces 0 and 1).
__device__ void fused_physbc_op(const GhostArgs& g, const FaceRel* __restrict__ frel, const GhostOp& op, int t) {
    const int mx = g.mx, my = g.my, mbc = g.mbc, meqn = g.meqn;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    double* qd = g.q + (long long) op.dst * g.pstride;
    auto code = [&](int f) { return (int) (signed char) ((op.bc >> (8 * f)) & 0xff); };
    const bool phys[4] = {code(0) == 1 || code(0) == 3, code(1) == 1 || code(1) == 3,
                          code(2) == 1 || code(2) == 3, code(3) == 1 || code(3) == 3};
    for (int face = 0; face < 4; face++) {
        if (!phys[face]) continue;
        const int mth = code(face);
        const int len = face < 2 ? my + 2 * mbc : mx + 2 * mbc;
        const int n = meqn * mbc * len;
        for (int e = t; e < n; e += 64) {
            const int m = e / (mbc * len), r = e - m * mbc * len;
            const int l = r / mbc, b = r - l * mbc + 1;
            int id, jd, is, js;
            if (face < 2) {
                jd = js = 1 - mbc + l;
                // the corners of two physical faces are the y-face's
                if ((jd < 1 && phys[2]) || (jd > my && phys[3])) continue;
                if (face == 0) { id = 1 - b; is = (mth == 1) ? 1 : b; }
                else { id = mx + b; is = (mth == 1) ? mx : mx + 1 - b; }
            } else {
                id = is = 1 - mbc + l;
                if (face == 2) { jd = 1 - b; js = (mth == 1) ? 1 : b; }
                else { jd = my + b; js = (mth == 1) ? my : my + 1 - b; }
            }
            const int mn = (face < 2) ? 1 : 2;  // normal momentum negated on walls
            const double x = fused_cell(g, frel, op.dst, is, js, m);
            qd[QIX(id, jd, m)] = (mth == 3 && m == mn) ? -x : x;
        }
    }
}

// Time-sync increments of one op (as time_sync_cells): a same-size op writes
// every position of its face, a coarse/fine op the positions its fine patch
// covers (the other child's op writes the rest; the reference adds zero for
// the positions a child does not cover).
__device__ void fused_sync_op(const FlatArgs& f, const int* __restrict__ op, int t) {
    const int mx = f.g.mx, meqn = f.g.meqn;
    const int type = op[0], dst = op[1], src = op[2], face = op[3], ti = op[4], tj = op[5];
    const double area = f.dxs[dst] * f.dys[dst];
    const double* __restrict__ rt = f.reg + (long long) dst * f.rstride;
    const double* __restrict__ rn = f.reg + (long long) src * f.rstride;
    double* __restrict__ qo = f.qinc + (long long) dst * f.cstride;
    const int kout = f.fluxreg ? REG_EFIN : REG_EFOUT;
    const int kn = face ^ 1;
    const int kfl = (face == 0 || face == 2) ? REG_FP : REG_FM;
    for (int e = t; e < meqn * mx; e += 64) {
        const int m = e / mx, c = e - m * mx + 1;
        auto R = [&](const double* rr, int kk, int kind, int pos) -> double {
            return rr[((kk * 4 + kind) * meqn + m) * mx + pos];
        };
        const double ein = R(rt, face, REG_EFIN, c - 1);
        const double ft = f.fluxreg ? 0.0 : R(rt, face, kfl, c - 1);
        double v;
        if (type != SYNC_F2C) {
            const double ef = R(rn, kn, kout, c - 1);
            const double fn = f.fluxreg ? 0.0 : R(rn, kn, kfl, c - 1);
            const double delta = (face == 0 || face == 2) ? (ef - ein) - (fn - ft) : (ein - ef) - (fn - ft);
            v = 0.5 * delta / area;
            // the partner's op (dst and src swapped, face ^ 1) would compute the
            // same delta: it reads these two registers and takes their
            // difference in the same order (the FP32 flux form stores no
            // fluctuations, fn = ft = 0)
            if (type == SYNC_SAMESIZE_PAIR) {
                const double area_src = f.dxs[src] * f.dys[src];
                f.qinc[(long long) src * f.cstride + (kn * meqn + m) * mx + c - 1] = 0.0 + 0.5 * delta / area_src;
            }
        } else {
            const int f0 = 2 * c - 1 + (face < 2 ? tj : ti);
            if (f0 < 1 || f0 + 1 > mx) continue;
            const double ef = R(rn, kn, kout, f0 - 1) + R(rn, kn, kout, f0);
            const double flu = f.fluxreg ? 0.0 : R(rn, kn, kfl, f0 - 1) + R(rn, kn, kfl, f0);
            const double deltac = (face == 0 || face == 2) ? (ef - ein) - (flu - ft) : (ein - ef) - (flu - ft);
            v = deltac / area;
        }
        qo[(face * meqn + m) * mx + c - 1] = 0.0 + v;
    }
}

// The fused launch, no op depending on another: ghost ops (copies and
// averages read interior cells; interpolations and physical BCs recompute the
// ghost cells they read), then time-sync increments.  A face, physical-BC or
// sync op takes a wave; the corner ops, a few cells each, take a quarter of
// one.  The interior cells keep their values before the time sync, which
// every ghost op reads (the reference fills ghost cells before it
// synchronises, src/fclaw2d_advance.c:330-341); the next update adds the
// increments to the cells it stages.
constexpr int FLAT_OPW = 4;           // waves per workgroup
constexpr int FLAT_CORNER_LANES = 16; // lanes of one corner op
constexpr int FLAT_CORNERS = 64 / FLAT_CORNER_LANES;

__device__ __forceinline__ int flat_corner_waves(int n) { return (n + FLAT_CORNERS - 1) / FLAT_CORNERS; }

__global__ void __launch_bounds__(64 * FLAT_OPW) ghost_flat(FlatArgs f) {
    int w = blockIdx.x * FLAT_OPW + (int) (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (w < f.nbig) {
        const GhostOp op = f.g.ops[f.ncorner0 + w];
        if (op.type == GHOST_INTERP_FACE || op.type == GHOST_INTERP_CORNER) fused_interp_op(f.g, f.frel, op, lane);
        else if (op.type == GHOST_PHYSBC) fused_physbc_op(f.g, f.frel, op, lane);
        else ghost_exchange_op(f.g, op, lane);
        return;
    }
    w -= f.nbig;
    const int nw0 = flat_corner_waves(f.ncorner0), nw1 = flat_corner_waves(f.ncorner1);
    if (w < nw0 + nw1) {
        // corner ops [0, ncorner0) and [ncorner0 + nbig, ... + ncorner1)
        const bool first = w < nw0;
        const int k = (first ? w : w - nw0) * FLAT_CORNERS + lane / FLAT_CORNER_LANES;
        if (k < (first ? f.ncorner0 : f.ncorner1)) {
            const GhostOp op = f.g.ops[(first ? 0 : f.ncorner0 + f.nbig) + k];
            const int t = lane % FLAT_CORNER_LANES;
            if (op.type == GHOST_INTERP_CORNER) fused_interp_op<FLAT_CORNER_LANES>(f.g, f.frel, op, t);
            else ghost_exchange_op<FLAT_CORNER_LANES>(f.g, op, t);
        }
        return;
    }
    w -= nw0 + nw1;
    if (w < f.nsync) fused_sync_op(f, f.sops + 8 * w, lane);
}

// The pending increments added into the slab (after the last step of an
// advance, so the state is the synchronised one), one thread per edge cell of
// the patches listed in slots (those with a corrected face), in the update's
// order: a corner cell belongs to its x-face when that face is
// corrected, and takes both faces' increments in the reference's order.
__global__ void apply_qcorr(double* __restrict__ q, const double* __restrict__ qcorr,
                            const int* __restrict__ cmask, const int* __restrict__ slots, int n,
                            long long pstride, long long cstride, int mx, int my, int mbc, int meqn) {
    const int per = 4 * meqn * mx;
    const long long e = (long long) blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long) n * per) return;
    const int k = (int) (e / per), r = (int) (e - (long long) k * per);
    const int slot = slots[k];
    const int fc = r / (meqn * mx), r2 = r - fc * meqn * mx, m = r2 / mx, c = r2 - m * mx + 1;
    const int mask = cmask[slot];
    if (!((mask >> fc) & 1)) return;
    const int len = fc < 2 ? my : mx;
    const int other = (c == 1 || c == len) ? (fc < 2 ? (c == 1 ? 2 : 3) : (c == 1 ? 0 : 1)) : -1;
    const bool other_on = other >= 0 && ((mask >> other) & 1);
    if (fc >= 2 && other_on) return;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    const int i = fc == 0 ? 1 : fc == 1 ? mx : c, j = fc == 2 ? 1 : fc == 3 ? my : c;
    const double* qc = qcorr + (long long) slot * cstride;
    const double v = qc[(fc * meqn + m) * len + c - 1];
    double* cell = q + (long long) slot * pstride + QIX(i, j, m);
    double x = *cell;
    if (other_on) {
        const double w = qc[(other * meqn + m) * mx + (fc == 0 ? 0 : mx - 1)];
        if (sync_y_first(mask, fc, other)) { x = x + w; x = x + v; }
        else { x = x + v; x = x + w; }
    } else {
        x = x + v;
    }
    *cell = x;
}

int launch_ghost_flat(const FlatArgs& f, hipStream_t st) {
    const int n = f.nbig + (f.ncorner0 + FLAT_CORNERS - 1) / FLAT_CORNERS +
                  (f.ncorner1 + FLAT_CORNERS - 1) / FLAT_CORNERS + f.nsync;
    if (n <= 0) return 0;
    hipLaunchKernelGGL(ghost_flat, dim3((n + FLAT_OPW - 1) / FLAT_OPW), dim3(64 * FLAT_OPW), 0, st, f);
    return (int) hipGetLastError();
}

int launch_apply_qcorr(double* q, const double* qcorr, const int* cmask, const int* slots, int n,
                       long long pstride, long long cstride, int mx, int my, int mbc, int meqn,
                       hipStream_t st) {
    const long long total = (long long) n * 4 * meqn * mx;
    if (total <= 0) return 0;
    hipLaunchKernelGGL(apply_qcorr, dim3((unsigned) ((total + 255) / 256)), dim3(256), 0, st, q, qcorr, cmask,
                       slots, n, pstride, cstride, mx, my, mbc, meqn);
    return (int) hipGetLastError();
}

__global__ void reset_scalar(unsigned long long* p) { *p = 0ull; }

// ---------------------------------------------------------------- regrid
// Gradient criterion of the reference's tag4refinement / tag4coarsening
// (src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_tag4refinement.f,
// src/patches/clawpatch/fort_user/fclaw2d_clawpatch_gradient_exceeds_th.f90,
// unmapped): does some interior cell have |(dq/dx, dq/dy, 0)| above the
// threshold?  out[2w] for the refine threshold, out[2w+1] for the coarsen
// threshold (a negative refine threshold or a coarsen threshold <= 0 always
// counts as exceeded: always refine, never coarsen, as
// src/patches/clawpatch/fclaw2d_clawpatch.cpp:813-817, 890-911 decide).
__global__ void __launch_bounds__(256)
tag_patches(const double* __restrict__ q, const int* __restrict__ slots, int n,
            const double* __restrict__ dxs, const double* __restrict__ dys, long long pstride,
            int mx, int my, int mbc, int ivar, double thr_refine, double thr_coarsen,
            int* __restrict__ out) {
    const int w = blockIdx.x;
    if (w >= n) return;
    const int slot = slots[w];
    const double* qp = q + (long long) slot * pstride;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    const double dx2 = 2 * dxs[slot], dy2 = 2 * dys[slot];
    int er = 0, ec = 0;
    for (int e = threadIdx.x; e < mx * my; e += blockDim.x) {
        const int j = e / mx + 1, i = e - (j - 1) * mx + 1;
        const double dqx = (qp[QIX(i + 1, j, ivar)] - qp[QIX(i - 1, j, ivar)]) / dx2;
        const double dqy = (qp[QIX(i, j + 1, ivar)] - qp[QIX(i, j - 1, ivar)]) / dy2;
        const double g3 = 0.0;
        const double ds = sqrt(dqx * dqx + dqy * dqy + g3 * g3);
        er |= ds > thr_refine;
        ec |= ds > thr_coarsen;
    }
    er = __syncthreads_or(er);
    ec = __syncthreads_or(ec);
    if (threadIdx.x == 0) {
        out[2 * w] = thr_refine < 0 ? 1 : er;
        out[2 * w + 1] = thr_coarsen <= 0 ? 1 : ec;
    }
}

// Repopulate one new patch (cb_fclaw2d_regrid_repopulate,
// src/fclaw2d_regrid.c:95-230): copy of the same old patch, interpolation from
// the old parent (fclaw2d_clawpatch46_fort_interpolate2fine, over the
// reference's loop ranges), or average of the four old children
// (fclaw2d_clawpatch46_fort_average2coarse).  op = (kind, dst, src0..src3, -, -).
__global__ void __launch_bounds__(256)
regrid_remap(const double* __restrict__ qold, double* __restrict__ qnew,
             const int* __restrict__ ops, int nops, long long pstride, int mx, int my, int mbc,
             int meqn) {
    const int o = blockIdx.x;
    if (o >= nops) return;
    const int* op = ops + 8 * o;
    const int kind = op[0];
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    double* qd = qnew + (long long) op[1] * pstride;
    if (kind == REGRID_SAME) {
        const double* qs = qold + (long long) op[2] * pstride;
        for (long long e = threadIdx.x; e < pstride; e += blockDim.x) qd[e] = qs[e];
    } else {
        // the cells the transfer does not write (ghost cells of new children
        // and parents) start at zero until the ghost fill
        for (long long e = threadIdx.x; e < pstride; e += blockDim.x) qd[e] = 0.0;
        __syncthreads();
    }
    if (kind == REGRID_CHILD) {
        const double* qc = qold + (long long) op[2] * pstride;
        const int igrid = op[3], ig = igrid % 2, jg = (igrid - ig) / 2;
        const int i1 = 1 - ig, ni = mx / 2 + 1, ic_add = ig * mx / 2;
        const int j1 = 1 - jg, nj = my / 2 + 1, jc_add = jg * my / 2;
        for (int e = threadIdx.x; e < meqn * ni * nj; e += blockDim.x) {
            const int m = e / (ni * nj), r = e - m * ni * nj;
            const int j = j1 + r / ni, i = i1 + r % ni;
            const int ic = i + ic_add, jc = j + jc_add;
            const double c = qc[QIX(ic, jc, m)];
            const double gradx = amr_slope(c - qc[QIX(ic - 1, jc, m)], qc[QIX(ic + 1, jc, m)] - c);
            const double grady = amr_slope(c - qc[QIX(ic, jc - 1, m)], qc[QIX(ic, jc + 1, m)] - c);
            for (int ii = 1; ii <= 2; ii++)
                for (int jj = 1; jj <= 2; jj++) {
                    const double shiftx = (ii - 2 / 2.0 - 0.5) / 2;
                    const double shifty = (jj - 2 / 2.0 - 0.5) / 2;
                    qd[QIX((i - 1) * 2 + ii, (j - 1) * 2 + jj, m)] = c + shiftx * gradx + shifty * grady;
                }
        }
    } else if (kind != REGRID_SAME) {
        for (int e = threadIdx.x; e < meqn * mx * my; e += blockDim.x) {
            const int m = e / (mx * my), r = e - m * mx * my;
            const int jc = r / mx + 1, ic = r - (jc - 1) * mx + 1;
            const int ig = ic > mx / 2, jg = jc > my / 2, g = ig + 2 * jg;
            const double* qf = qold + (long long) op[2 + g] * pstride;
            const int i = ic - ig * mx / 2, j = jc - jg * mx / 2;  // mx, as the reference writes it
            double sum = 0;
            sum = sum + qf[QIX(2 * i - 1, 2 * j - 1, m)];
            sum = sum + qf[QIX(2 * i, 2 * j - 1, m)];
            sum = sum + qf[QIX(2 * i - 1, 2 * j, m)];
            sum = sum + qf[QIX(2 * i, 2 * j, m)];
            qd[QIX(ic, jc, m)] = sum / 4;
        }
    }
}

__global__ void copy_patches(const double* __restrict__ src, double* __restrict__ dst,
                             const int* __restrict__ slots, int n, long long pstride) {
    const int w = blockIdx.y;
    if (w >= n) return;
    const long long base = (long long) slots[w] * pstride;
    for (long long e = blockIdx.x * (long long) blockDim.x + threadIdx.x; e < pstride;
         e += (long long) gridDim.x * blockDim.x)
        dst[base + e] = src[base + e];
}

// fclaw2d_clawpatch46_fort_timeinterp
// (src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_timeinterp.f:1-58):
// qinterp = qlast + alpha (qcurr - qlast) on the interior cells within mbc of
// the patch edge (the cells fine neighbours interpolate from).
__global__ void timeinterp_band(const double* __restrict__ q, const double* __restrict__ ql,
                                double* __restrict__ qi, const int* __restrict__ slots, int n,
                                long long pstride, int mx, int my, int mbc, int meqn,
                                double alpha) {
    const int w = blockIdx.y;
    if (w >= n) return;
    const long long base = (long long) slots[w] * pstride;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < meqn * mx * my; e += gridDim.x * blockDim.x) {
        const int m = e / (mx * my), r = e - m * mx * my;
        const int j = r / mx + 1, i = r - (j - 1) * mx + 1;
        if (i > mbc && i <= mx - mbc && j > mbc && j <= my - mbc) continue;  // the hole
        const long long o = base + QIX(i, j, m);
        qi[o] = ql[o] + alpha * (q[o] - ql[o]);
    }
}

__global__ void reset_registers(double* __restrict__ reg, const int* __restrict__ resets,
                                int nresets, long long rstride, int per_face) {
    const int r = blockIdx.x;
    if (r >= nresets) return;
    double* p = reg + (long long) resets[2 * r] * rstride + (long long) resets[2 * r + 1] * per_face;
    for (int e = threadIdx.x; e < per_face; e += blockDim.x) p[e] = 0.0;
}

__global__ void copy_interior(const double* __restrict__ src, double* __restrict__ dst,
                              const int* __restrict__ slots, int n, long long pstride,
                              int mx, int my, int mbc, int meqn) {
    const int w = blockIdx.y;
    if (w >= n) return;
    const long long base = (long long) slots[w] * pstride;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < meqn * mx * my;
         e += gridDim.x * blockDim.x) {
        const int m = e / (mx * my), r = e - m * mx * my;
        const int j = r / mx + 1, i = r - (j - 1) * mx + 1;
        dst[base + QIX(i, j, m)] = src[base + QIX(i, j, m)];
    }
}

// Error-free transformation a + b = s + e (Knuth TwoSum).
__device__ __forceinline__ void two_sum(double a, double b, double& s, double& e) {
    s = a + b;
    const double bb = s - a;
    e = (a - (s - bb)) + (b - bb);
}

// Compensated per-patch sums of q*dx*dy: each thread keeps a double-double
// accumulator, then the workgroup combines them with TwoSum.
__global__ void __launch_bounds__(256)
patch_mass(const double* __restrict__ q, const int* __restrict__ slots, int n,
           const double* __restrict__ dxs, const double* __restrict__ dys, long long pstride,
           int mx, int my, int mbc, int meqn, double* __restrict__ out) {
    __shared__ double sh[2][256];
    const int w = blockIdx.x;
    if (w >= n) return;
    const int slot = slots[w];
    const double area = dxs[slot] * dys[slot];
    const double* qp = q + (long long) slot * pstride;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    for (int m = 0; m < meqn; m++) {
        double hi = 0.0, lo = 0.0;
        for (int e = threadIdx.x; e < mx * my; e += 256) {
            const int j = e / mx + 1, i = e - (j - 1) * mx + 1;
            double s, err;
            two_sum(hi, qp[QIX(i, j, m)] * area, s, err);
            hi = s;
            lo += err;
        }
        sh[0][threadIdx.x] = hi;
        sh[1][threadIdx.x] = lo;
        __syncthreads();
        for (int st = 128; st >= 1; st >>= 1) {
            if ((int) threadIdx.x < st) {
                double s, err;
                two_sum(sh[0][threadIdx.x], sh[0][threadIdx.x + st], s, err);
                sh[0][threadIdx.x] = s;
                sh[1][threadIdx.x] = sh[1][threadIdx.x] + sh[1][threadIdx.x + st] + err;
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            out[((long long) slot * meqn + m) * 2 + 0] = sh[0][0];
            out[((long long) slot * meqn + m) * 2 + 1] = sh[1][0];
        }
        __syncthreads();
    }
}

int launch_ghost_exchange(const GhostArgs& g, hipStream_t st) {
    if (g.nops <= 0) return 0;
    hipLaunchKernelGGL(ghost_phase<4>, dim3((g.nops + 3) / 4), dim3(256), 0, st, g);
    return (int) hipGetLastError();
}

int launch_ghost_groups(const GhostArgs& g, const int4* groups, int ngroups, hipStream_t st) {
    if (ngroups <= 0) return 0;
    hipLaunchKernelGGL(ghost_groups, dim3(ngroups), dim3(PHYSBC_THREADS), 0, st, g, groups, ngroups);
    return (int) hipGetLastError();
}

int launch_ghost_physbc(const GhostArgs& g, hipStream_t st) {
    if (g.nops <= 0) return 0;
    hipLaunchKernelGGL(physbc_phase, dim3(g.nops), dim3(PHYSBC_THREADS), 0, st, g);
    return (int) hipGetLastError();
}

int launch_copy_patches(const double* src, double* dst, const int* slots, int n, long long pstride,
                        hipStream_t st) {
    if (n <= 0) return 0;
    dim3 grid((unsigned) ((pstride + 255) / 256), n);
    hipLaunchKernelGGL(copy_patches, grid, dim3(256), 0, st, src, dst, slots, n, pstride);
    return (int) hipGetLastError();
}

// Ghost-patch message of each listed slot (the data a rank sends another for
// a patch bordering it): the q band of mbc ghost and refratio*mbc interior
// layers in the order of fclaw2d_clawpatch46_fort_local_ghost_pack
// (src/patches/clawpatch/fort_4.6/fclaw2d_clawpatch46_local_ghost_pack.f:38-84;
// band[] holds the slab offsets of one field in that order), then, with time
// sync, the registers in the order of fclaw2d_clawpatch_time_sync_pack_registers
// (src/patches/clawpatch/fclaw2d_clawpatch_conservation.c:112-195): per face
// meqn*len quadruples (fm, fp, edge flux 0, edge flux 1), then len pairs
// (edge length, area).  unpack = 1 writes q and the registers back (edge
// lengths and areas follow from dx, dy and are not read).
__global__ void __launch_bounds__(256)
ghost_message(double* __restrict__ q, double* __restrict__ reg, const int* __restrict__ slots,
              int n, const int* __restrict__ band, int nband, long long pstride, long long rstride,
              int mx, int my, int mbc, int meqn, const double* __restrict__ dxs,
              const double* __restrict__ dys, double* __restrict__ buf, long long elems,
              int unpack) {
    const int p = blockIdx.y;
    if (p >= n) return;
    const int slot = slots[p];
    double* qp = q + (long long) slot * pstride;
    double* b = buf + (long long) p * elems;
    const long long plane = (long long) (mx + 2 * mbc) * (my + 2 * mbc);
    const int nq = meqn * nband;
    const int stride = gridDim.x * blockDim.x;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nq; k += stride) {
        const int m = k / nband, e = k - m * nband;
        double* c = qp + m * plane + band[e];
        if (unpack) *c = b[k];
        else b[k] = *c;
    }
    if (!reg) return;
    double* rp = reg + (long long) slot * rstride;
    double* rb = b + nq;
    const double dx = dxs[slot], dy = dys[slot];
    const int nreg = 2 * (4 * meqn + 2) * (mx + my);
    for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nreg; r += stride) {
        int f = 0, t = r;
        for (; f < 3; f++) {
            const int len = f < 2 ? my : mx, blk = (4 * meqn + 2) * len;
            if (t < blk) break;
            t -= blk;
        }
        const int len = f < 2 ? my : mx;
        if (t < 4 * meqn * len) {
            const int j = t >> 2, kind = t & 3;        // kinds FM, FP, EFIN, EFOUT in order
            const int m = j / len, pos = j - m * len;
            double* c = rp + ((f * 4 + kind) * meqn + m) * mx + pos;
            if (unpack) *c = rb[r];
            else rb[r] = *c;
        } else if (!unpack) {
            const int w = (t - 4 * meqn * len) & 1;
            rb[r] = w == 0 ? (f < 2 ? dy : dx) : dx * dy;
        }
    }
}

int launch_ghost_message(double* q, double* reg, const int* slots, int n, const int* band,
                         int nband, long long pstride, long long rstride, int mx, int my, int mbc,
                         int meqn, const double* dxs, const double* dys, double* buf,
                         long long elems, int unpack, hipStream_t st) {
    if (n <= 0) return 0;
    const long long per = elems;
    dim3 grid((unsigned) std::min<long long>((per + 255) / 256, 64), n);
    hipLaunchKernelGGL(ghost_message, grid, dim3(256), 0, st, q, reg, slots, n, band, nband,
                       pstride, rstride, mx, my, mbc, meqn, dxs, dys, buf, elems, unpack);
    return (int) hipGetLastError();
}

// Cell values of the VTK "meqn" array: per patch, cells j = 1..my, i = 1..mx,
// then the meqn fields of the cell, as float (fclaw2d_output_vtk_value_cb,
// src/patches/clawpatch/fclaw2d_clawpatch_output_vtk.c:781-808).
__global__ void __launch_bounds__(256)
vtk_cell_values(const double* __restrict__ q, int first, long long pstride, int mx, int my,
                int mbc, int meqn, float* __restrict__ out) {
    const int p = blockIdx.y;
    const int per = mx * my * meqn;
    const double* qp = q + (long long) (first + p) * pstride;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < per; e += gridDim.x * blockDim.x) {
        const int cell = e / meqn, k = e - cell * meqn;
        const int j = cell / mx, i = cell - j * mx;
        out[(long long) p * per + e] = (float) qp[((long long) k * nyp + j + mbc) * nxp + i + mbc];
    }
}

int launch_vtk_cell_values(const double* q, int first, int count, long long pstride, int mx, int my,
                           int mbc, int meqn, float* out, hipStream_t st) {
    if (count <= 0) return 0;
    const int per = mx * my * meqn;
    dim3 grid((unsigned) std::min(64, (per + 255) / 256), count);
    hipLaunchKernelGGL(vtk_cell_values, grid, dim3(256), 0, st, q, first, pstride, mx, my, mbc,
                       meqn, out);
    return (int) hipGetLastError();
}

int launch_timeinterp(const double* q, const double* qlast, double* qti, const int* slots, int n,
                      long long pstride, int mx, int my, int mbc, int meqn, double alpha,
                      hipStream_t st) {
    if (n <= 0) return 0;
    dim3 grid((meqn * mx * my + 255) / 256, n);
    hipLaunchKernelGGL(timeinterp_band, grid, dim3(256), 0, st, q, qlast, qti, slots, n, pstride,
                       mx, my, mbc, meqn, alpha);
    return (int) hipGetLastError();
}

int launch_tag_patches(const double* q, const int* slots, int n, const double* dxs,
                       const double* dys, long long pstride, int mx, int my, int mbc, int ivar,
                       double thr_refine, double thr_coarsen, int* out, hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(tag_patches, dim3(n), dim3(256), 0, st, q, slots, n, dxs, dys, pstride, mx,
                       my, mbc, ivar, thr_refine, thr_coarsen, out);
    return (int) hipGetLastError();
}

int launch_regrid_remap(const double* qold, double* qnew, const int* ops, int nops,
                        long long pstride, int mx, int my, int mbc, int meqn, hipStream_t st) {
    if (nops <= 0) return 0;
    hipLaunchKernelGGL(regrid_remap, dim3(nops), dim3(256), 0, st, qold, qnew, ops, nops, pstride,
                       mx, my, mbc, meqn);
    return (int) hipGetLastError();
}

int launch_reset_scalar(unsigned long long* p, hipStream_t st) {
    hipLaunchKernelGGL(reset_scalar, dim3(1), dim3(1), 0, st, p);
    return (int) hipGetLastError();
}

int launch_copy_interior(const double* src, double* dst, const int* slots, int n,
                         long long pstride, int mx, int my, int mbc, int meqn,
                         hipStream_t st) {
    if (n <= 0) return 0;
    const int per = meqn * mx * my;
    dim3 grid((per + 255) / 256, n);
    hipLaunchKernelGGL(copy_interior, grid, dim3(256), 0, st, src, dst, slots, n, pstride, mx,
                       my, mbc, meqn);
    return (int) hipGetLastError();
}

int launch_mass(const double* q, const int* slots, int n, const double* dxs, const double* dys,
                long long pstride, int mx, int my, int mbc, int meqn, double* out,
                hipStream_t st) {
    if (n <= 0) return 0;
    hipLaunchKernelGGL(patch_mass, dim3(n), dim3(256), 0, st, q, slots, n, dxs, dys, pstride,
                       mx, my, mbc, meqn, out);
    return (int) hipGetLastError();
}

// ---------------------------------------------------------------- time sync
// The reference's fclaw2d_time_sync (src/fclaw2d_time_sync.c:118-202) with its
// Fortran kernels clawpack46_fort_time_sync_samesize / _f2c
// (src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_time_sync.f:249-700):
// the coarse patch's edge flux is replaced by the sum of the fine fluxes
// across each coarse/fine face, and half the mismatch between same-size
// neighbours goes to each side.  One workgroup per destination patch runs its
// ops in order (same-size faces 0..3, then coarse/fine faces), so a corner
// cell receives its corrections in the reference's order.  Registers are
// read only; every op writes cells of its own destination patch.
__global__ void __launch_bounds__(128) time_sync_groups(SyncArgs a) {
    const int g = blockIdx.x;
    if (g >= a.ngroups) return;
    const int mx = a.mx, mbc = a.mbc, meqn = a.meqn, nx = mx + 2 * mbc;
    const int kout = a.fluxreg ? REG_EFIN : REG_EFOUT;
    for (int o = a.start[g]; o < a.start[g + 1]; o++) {
        const int* op = a.ops + 8 * o;
        const int type = op[0], dst = op[1], src = op[2], face = op[3], ti = op[4], tj = op[5];
        double* __restrict__ q = a.q + (long long) dst * a.pstride;
        const double* __restrict__ rt = a.reg + (long long) dst * a.rstride;
        const double* __restrict__ rn = a.reg + (long long) src * a.rstride;
        const double area = a.dxs[dst] * a.dys[dst];
        const int kn = face ^ 1;              // the neighbour's face on this interface
        const int kfl = (face == 0 || face == 2) ? REG_FP : REG_FM;
        for (int e = threadIdx.x; e < meqn * mx; e += blockDim.x) {
            const int m = e / mx, c = e - m * mx + 1;   // c = 1..mx along the face
            auto R = [&](const double* r, int k, int kind, int pos) -> double {
                return r[((k * 4 + kind) * meqn + m) * mx + pos];
            };
            int i, j;
            if (face == 0) { i = 1; j = c; }
            else if (face == 1) { i = mx; j = c; }
            else if (face == 2) { i = c; j = 1; }
            else { i = c; j = mx; }
            double* qc = q + ((long long) m * nx + (j + mbc - 1)) * nx + (i + mbc - 1);
            const double ein = R(rt, face, REG_EFIN, c - 1);
            const double ft = a.fluxreg ? 0.0 : R(rt, face, kfl, c - 1);
            if (type == SYNC_SAMESIZE) {
                const double ef = R(rn, kn, kout, c - 1);
                const double fn = a.fluxreg ? 0.0 : R(rn, kn, kfl, c - 1);
                const double delta = (face == 0 || face == 2) ? (ef - ein) - (fn - ft)
                                                              : (ein - ef) - (fn - ft);
                *qc = *qc + 0.5 * delta / area;
            } else {
                const int f0 = 2 * c - 1 + (face < 2 ? tj : ti);  // fine rows/columns f0, f0+1
                double deltac = 0.0;
                if (f0 >= 1 && f0 + 1 <= mx) {
                    const double ef = R(rn, kn, kout, f0 - 1) + R(rn, kn, kout, f0);
                    const double flu = a.fluxreg ? 0.0 : R(rn, kn, kfl, f0 - 1) + R(rn, kn, kfl, f0);
                    deltac = (face == 0 || face == 2) ? (ef - ein) - (flu - ft) : (ein - ef) - (flu - ft);
                }
                *qc = *qc + deltac / area;
            }
        }
        __syncthreads();
    }
}

// The same corrections in two passes: every (op, field, position) of the
// group computes its increment into LDS at once, then every edge cell adds
// the increments of the ops on its face(s) in op order and is written once.
// A corner cell belongs to the x-face thread, which also takes the y-face
// ops' increments at the matching position; the sums are the one-op-at-a-time
// sums, operation for operation.
__global__ void __launch_bounds__(256) time_sync_cells(SyncArgs a) {
    extern __shared__ double inc[];   // [maxgroup][meqn][mx], then the ops' faces
    const int g = blockIdx.x;
    if (g >= a.ngroups) return;
    const int mx = a.mx, mbc = a.mbc, meqn = a.meqn, nx = mx + 2 * mbc;
    const int per = meqn * mx;
    const int o0 = a.start[g], nop = a.start[g + 1] - o0;
    int* oface = (int*) (inc + (long long) a.maxgroup * per);
    const int kout = a.fluxreg ? REG_EFIN : REG_EFOUT;
    const int dst = a.ops[8 * o0 + 1];
    const double area = a.dxs[dst] * a.dys[dst];
    const double* __restrict__ rt = a.reg + (long long) dst * a.rstride;
    for (int k = threadIdx.x; k < nop; k += blockDim.x) oface[k] = a.ops[8 * (o0 + k) + 3];
    for (int e = threadIdx.x; e < nop * per; e += blockDim.x) {
        const int k = e / per, r = e - k * per, m = r / mx, c = r - m * mx + 1;
        const int* op = a.ops + 8 * (o0 + k);
        const int type = op[0], src = op[2], face = op[3], ti = op[4], tj = op[5];
        const double* __restrict__ rn = a.reg + (long long) src * a.rstride;
        const int kn = face ^ 1;
        const int kfl = (face == 0 || face == 2) ? REG_FP : REG_FM;
        auto R = [&](const double* rr, int kk, int kind, int pos) -> double {
            return rr[((kk * 4 + kind) * meqn + m) * mx + pos];
        };
        const double ein = R(rt, face, REG_EFIN, c - 1);
        const double ft = a.fluxreg ? 0.0 : R(rt, face, kfl, c - 1);
        double v;
        if (type == SYNC_SAMESIZE) {
            const double ef = R(rn, kn, kout, c - 1);
            const double fn = a.fluxreg ? 0.0 : R(rn, kn, kfl, c - 1);
            const double delta = (face == 0 || face == 2) ? (ef - ein) - (fn - ft) : (ein - ef) - (fn - ft);
            v = 0.5 * delta / area;
        } else {
            const int f0 = 2 * c - 1 + (face < 2 ? tj : ti);
            double deltac = 0.0;
            if (f0 >= 1 && f0 + 1 <= mx) {
                const double ef = R(rn, kn, kout, f0 - 1) + R(rn, kn, kout, f0);
                const double flu = a.fluxreg ? 0.0 : R(rn, kn, kfl, f0 - 1) + R(rn, kn, kfl, f0);
                deltac = (face == 0 || face == 2) ? (ef - ein) - (flu - ft) : (ein - ef) - (flu - ft);
            }
            v = deltac / area;
        }
        inc[e] = v;
    }
    __syncthreads();
    double* __restrict__ q = a.q + (long long) dst * a.pstride;
    for (int e = threadIdx.x; e < 4 * per; e += blockDim.x) {
        const int f = e / per, r = e - f * per, m = r / mx, c = r - m * mx + 1;
        if (f >= 2 && (c == 1 || c == mx)) continue;   // corners: the x-face thread
        int i, j;
        if (f == 0) { i = 1; j = c; }
        else if (f == 1) { i = mx; j = c; }
        else if (f == 2) { i = c; j = 1; }
        else { i = c; j = mx; }
        int f2 = -1, c2 = 0;
        if (f < 2 && (c == 1 || c == mx)) { f2 = c == 1 ? 2 : 3; c2 = f == 0 ? 1 : mx; }
        double* qc = q + ((long long) m * nx + (j + mbc - 1)) * nx + (i + mbc - 1);
        double v = 0.0;
        bool touched = false;
        for (int k = 0; k < nop; k++) {
            const int fk = oface[k];
            if (fk != f && fk != f2) continue;
            const double d = inc[(k * meqn + m) * mx + (fk == f ? c : c2) - 1];
            v = touched ? v + d : *qc + d;
            touched = true;
        }
        if (touched) *qc = v;
    }
}

int launch_time_sync(const SyncArgs& a, hipStream_t st) {
    if (a.ngroups > 0) {
        const size_t lds = sizeof(double) * (size_t) a.maxgroup * a.meqn * a.mx + sizeof(int) * (size_t) a.maxgroup;
        if (lds <= 32768)
            hipLaunchKernelGGL(time_sync_cells, dim3(a.ngroups), dim3(256), lds, st, a);
        else
            hipLaunchKernelGGL(time_sync_groups, dim3(a.ngroups), dim3(128), 0, st, a);
        if (hipError_t e = hipGetLastError()) return (int) e;
    }
    if (a.nresets > 0) {
        // after every correction of this synchronisation (they read the registers)
        const int per_face = 4 * a.meqn * a.mx;
        hipLaunchKernelGGL(reset_registers, dim3(a.nresets), dim3(128), 0, st,
                           const_cast<double*>(a.reg), a.resets, a.nresets, a.rstride, per_face);
    }
    return (int) hipGetLastError();
}

// ---------------------------------------------------------------- FP64 step
int launch_step_fp64(int solver, int tile, int waves, const StepArgs& a, const StepLaunch& st) {
    switch (solver) {
    case 0: return launch_step_tile<RPAdvection<double>, double, true>(tile, waves, a, st);
    case 1: return launch_step_tile<RPShallow<double>, double, true, true>(tile, waves, a, st);
    case 2: return launch_step_tile<RPEuler<double>, double, true>(tile, waves, a, st);
    default: return (int) hipErrorInvalidValue;
    }
}

int default_waves_fp64(int /*solver*/, int tile) { return tile == 16 ? default_waves_t16<double>() : 0; }

}  // namespace hipclaw
