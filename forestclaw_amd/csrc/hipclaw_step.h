// Batched patch-update kernel.  A persistent workgroup loops over the T x T
// tiles of every patch in a work list (all levels of the forest in one
// launch) and updates each tile's interior cells.
//
// What it computes is the reference's Clawpack-4.6 step
// (src/solvers/fc2d_clawpack4.6/fortran_source/clawpack46_step2.f:5-239,
//  clawpack46_flux2.f:4-228, clawpack46_inlinelimiter.f:39-115,
//  clawpack46_step2_wrap.f:106-125) restricted to the interior cells of the
// tile.  How it computes it is MI355X-first:
//
//  * the tile and its 2-cell halo are staged once in LDS (FP64);
//  * while a tile is being computed, the next tile's cells are already in
//    flight from HBM into registers (software pipelining across tiles), so
//    the global-load latency is hidden behind the sweeps;
//  * each lane owns ONE staged cell and the interface on its left: a row (or
//    column) of the T+4 staged cells lives in T+4 consecutive lanes of a
//    wave64, so the left cell's values, the limiter's neighbour waves and the
//    transverse fluctuations of the next interface are exchanged with
//    cross-lane DPP moves, not memory;
//  * no global scratch: normal fluxes and transverse corrections are
//    accumulated in LDS in exactly the order the reference adds them, so the
//    FP64 instantiation (compiled with FMA contraction off) is bit-identical to
//    the CPU Fortran;
//  * the CFL number is reduced in registers over all tiles of a workgroup and
//    with wave64 shuffles, then one FP64 atomic max per wave.
//
// FLUXFORM selects the FP32-bulk variant for conservative solvers: the Riemann
// solves, limiter and transverse splitting run in R = float, but the update is
// written in conservative flux form with FP64 interface fluxes
// F = f(q_left) + fm shared by both neighbours, so mass is conserved to FP64
// rounding.
#pragma once

#include "hipclaw_rp.h"

namespace hipclaw {

struct StepArgs {
    const double* qsrc;        // [slot][meqn][my+2mbc][mx+2mbc]
    double* qdst;              // may equal qsrc when one tile covers a patch
    const double* aux;         // [slot][maux][my+2mbc][mx+2mbc] or null
    const int* work;           // patch slots to update
    const double* dxs;         // [slot]
    const double* dys;         // [slot]
    unsigned long long* cfl;   // FP64 bits, atomic max (non-negative doubles order as integers)
    long long pstride;         // doubles per patch in q
    long long astride;         // doubles per patch in aux
    int nwork;
    int mx, my, mbc;
    int ntx, nty;              // tiles per patch in x and y
    int order2;                // method(2) == 2
    int transverse;            // method(3)
    int mthlim[10];            // limiter per wave family (mwaves <= 10)
    double dt;
    PhysParams pp;
    // Time-sync flux registers (null: not kept).  Per slot
    // reg[((face * 4 + kind) * meqn + m) * mx + pos], face 0..3 = left, right,
    // bottom, top; pos = j-1 (faces 0/1) or i-1 (faces 2/3).  The FP64 form
    // writes the step's edge fluctuations (kinds REG_FM, REG_FP) and the
    // conservative flux at the first interior / first ghost cell (REG_EFIN,
    // REG_EFOUT); the FP32 flux form writes its shared FP64 edge flux in REG_EFIN.
    double* reg;
    long long rstride;
    int reg_accumulate;   // add to the registers (subcycling) instead of overwriting them
    const double* dtscale;  // per-slot factor on dt (an ensemble of runs), or null
    // Deferred conservation fix-up (null: none pending).  qcorr[slot][face][m][pos]
    // holds the time-sync increment of interior edge cell pos of `face` for the
    // faces set in bits 0-3 of cmask[slot] (bits 4-7: that face's increment
    // comes from a coarse/fine op); qsrc still holds the values before the
    // fix-up, and staging adds the increments (sync_corner_order).
    const double* qcorr = nullptr;
    const int* cmask = nullptr;
    long long cstride = 0;
    // Ghost strips staged straight from a same-size neighbour (null: every halo
    // cell from the patch's own slab).  pull[slot][8]: for faces 0..3, then
    // corners 0..3, the offset in doubles from a ghost cell of `slot` to the
    // neighbour's interior cell it copies (fclaw2d_clawpatch46_copy.f), 0 where
    // the ghost cells come from the patch's own slab.  Only with two slabs.
    const int* pull = nullptr;
};

// The reference applies a patch's same-size corrections (faces 0..3) before
// its coarse/fine ones (src/fclaw2d_time_sync.c:126-151), so a corner cell
// takes its y-face increment first only when its x-face is coarse/fine and its
// y-face same-size.
__host__ __device__ __forceinline__ bool sync_y_first(int mask, int fx, int fy) {
    return ((mask >> (4 + fx)) & 1) && !((mask >> (4 + fy)) & 1);
}

enum { REG_FM = 0, REG_FP = 1, REG_EFIN = 2, REG_EFOUT = 3 };

// Register update r = r + v (r = r - v): the reference accumulates onto
// registers that start at zero (clawpack46_time_sync.f:75-147, 192-209); when
// every register is reset at every synchronisation (no subcycling) the
// previous value is zero and need not be read.
__device__ __forceinline__ void reg_add(double* r, double v, int accumulate) {
    *r = (accumulate ? *r : 0.0) + v;
}
__device__ __forceinline__ void reg_sub(double* r, double v, int accumulate) {
    *r = (accumulate ? *r : 0.0) - v;
}

template <typename R>
__device__ __forceinline__ R philim(int mth, R r) {
    switch (mth) {
    case 2: return rmax(rmax(R(0), rmin(R(1), R(2) * r)), rmin(R(2), r));
    case 3: return dv(r + fabs(r), R(1) + fabs(r));
    case 4: {
        const R c = dv(R(1) + r, R(2));
        return rmax(R(0), rmin(rmin(c, R(2)), R(2) * r));
    }
    case 5: return r;
    case 6: {
        const R th = (R) 1.3f;  // REAL*4 literal in clawpack46_inlinelimiter.f:101
        return rmax(R(0), rmin(rmin(th * r, dv(R(1) + r, R(2))), th));
    }
    default: return rmax(R(0), rmin(R(1), r));  // 1 (and the computed-GOTO fall-through)
    }
}

// Neighbour-lane exchange with DPP wavefront shifts (one VALU move per 32
// bits, no LDS round trip): prev_lane(v) in lane l is v of lane l-1
// (wave_shr:1), next_lane(v) is v of lane l+1 (wave_shl:1).  The lanes at the
// wave's ends read 0; no caller uses those values.
template <int CTRL>
__device__ __forceinline__ float dpp_move(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const unsigned long long b = (unsigned long long) __double_as_longlong(v);
    const unsigned lo = (unsigned) __builtin_amdgcn_update_dpp(0, (int) (unsigned) b, CTRL, 0xF, 0xF, true);
    const unsigned hi = (unsigned) __builtin_amdgcn_update_dpp(0, (int) (unsigned) (b >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long) (((unsigned long long) hi << 32) | lo));
}
constexpr int DPP_WAVE_SHL1 = 0x130;
constexpr int DPP_WAVE_SHR1 = 0x138;
template <typename R>
__device__ __forceinline__ R prev_lane(R v) { return dpp_move<DPP_WAVE_SHR1>(v); }
template <typename R>
__device__ __forceinline__ R next_lane(R v) { return dpp_move<DPP_WAVE_SHL1>(v); }

// The reference skips the limiter for a wave of zero norm
// (clawpack46_inlinelimiter.f:51).  FP64 keeps that exact test; in FP32 a norm
// in the denormal range is treated as zero, since the hardware reciprocal
// flushes it and dotl/wn would become 0 * inf.
__device__ __forceinline__ bool wave_nonzero(double wn) { return wn != 0.0; }
__device__ __forceinline__ bool wave_nonzero(float wn) { return wn >= 1.17549435e-38f; }

// LDS layout.  Rows that a wave walks across lanes in the transverse direction
// get an odd stride so that the 64 lanes of a y sweep (or of the final stage)
// fall into distinct banks.
template <int T, int MEQN, int MAUX, typename R, bool FLUXFORM, bool REGS = false>
struct TileSmem {
    static constexpr int QW = T + 4;            // q tile width (cells i0-2 .. i0+T+1)
    static constexpr int QS = QW | 1;           // padded row stride
    static constexpr int QP = QW * QS;          // one field
    static constexpr int GS = T + 1;            // padded row stride of gt and ab
    static constexpr int NFM = MEQN * T * (T + 1);          // x-interface slot s = i - i0 in [0, T], row b
    static constexpr int NFP = FLUXFORM ? 0 : NFM;
    static constexpr int NGT = MEQN * (T + 1) * GS;         // y-interface slot t = j - j0 in [0, T], col a
    static constexpr int NAB = 2 * MEQN * (T + 2) * GS;     // transverse corrections (x sweep, then y sweep)
    static constexpr size_t WORK_BYTES = sizeof(R) * (size_t) (NFM + NFP + NGT + NAB);
    double q[MEQN * QP];
    R aux[MAUX > 0 ? MAUX * QP : 1];
    double work[(WORK_BYTES + 7) / 8];
    // gm, gp at the patch's bottom and top edges, kept for the time-sync registers
    R eg[(REGS && !FLUXFORM) ? 2 * 2 * MEQN * T : 1];
    __device__ R* fm() { return reinterpret_cast<R*>(work); }
    __device__ R* fp() { return fm() + NFM; }
    __device__ R* gt() { return fp() + NFP; }
    __device__ R* ab() { return gt() + NGT; }
};

template <class RP, typename R, int T, bool FLUXFORM>
constexpr size_t tile_smem_bytes() {
    return sizeof(TileSmem<T, RP::MEQN, RP::MAUX, R, FLUXFORM>);
}

// Per-cell values a Riemann functor can take from the step kernel
// (RP::NCELL, RP::cell, RP::rpn2c), 0 if it has none.
template <class RP, class = void>
struct cell_values {
    static constexpr int N = 0;
};
template <class RP>
struct cell_values<RP, decltype((void) RP::NCELL)> {
    static constexpr int N = RP::NCELL;
};

// One interface's Riemann solve.  With per-cell values, each lane evaluates
// its own (right) cell's and takes the left cell's from the lane before: the
// kernel lays a row's cells on consecutive lanes, so every cell is evaluated
// once per sweep instead of once per side of each interface.
template <class RP, int IXY, int NC, typename R>
__device__ __forceinline__ void riemann(const R* qL, const R* qR, const R* auxL, const R* auxR,
                                        const PhysParams& pp, R wave[RP::MWAVES][RP::MEQN],
                                        R s[RP::MWAVES], R amdq[RP::MEQN], R apdq[RP::MEQN],
                                        R roe[RP::NROE]) {
    if constexpr (NC > 0) {
        R cL[NC], cR[NC];
        RP::template cell<IXY>(qR, pp, cR);
#pragma unroll
        for (int i = 0; i < NC; i++) cL[i] = prev_lane(cR[i]);
        RP::template rpn2c<IXY>(qL, qR, cL, cR, auxL, auxR, pp, wave, s, amdq, apdq, roe);
    } else {
        RP::template rpn2<IXY>(qL, qR, auxL, auxR, pp, wave, s, amdq, apdq, roe);
    }
}

// Waves per SIMD the register allocation aims for: 4, except FP64 16 x 16
// tiles, which need more than the 128 VGPRs of 4 waves and are held to the 168
// of 3 (without the target the allocator let them grow to 2 waves).
//
// LIM = 0 reads each wave family's limiter from a.mthlim at run time; LIM > 0
// is that limiter for every family, fixed at compile time.  The run-time
// switch keeps a per-family limiter code in scalar registers through both
// sweeps and, with the kernel's other uniform values, spills scalar
// registers into vector lanes; the fixed form has neither (about 40 % fewer
// vector instructions in each sweep for the MC limiter).
//
// A solver with many equations and waves (more than 16 wave components per
// interface) cannot keep them in the registers of 4 waves; it gets no target
// and the allocator sizes the kernel itself.
template <class RP, typename R, int T>
constexpr int step_waves_per_eu() {
    if (RP::MEQN * RP::MWAVES > 16) return 1;
#ifdef HIPCLAW_WPE_F32_T16
    if (sizeof(R) == 4 && T == 16) return HIPCLAW_WPE_F32_T16;
#endif
    return (sizeof(R) == 8 && T == 16) ? 3 : 4;
}

template <class RP, typename R, int T, int NW, bool FLUXFORM, bool REGS, int LIM = 0>
__global__ void __launch_bounds__(NW * 64) __attribute__((amdgpu_waves_per_eu(step_waves_per_eu<RP, R, T>())))
step_tiles(StepArgs a) {
    constexpr int MEQN = RP::MEQN, MW = RP::MWAVES, MAUX = RP::MAUX, NROE = RP::NROE;
    // wave components zero by construction, skipped in the sums over waves
    using WZ1 = wave_zeros<RP, 1>;
    using WZ2 = wave_zeros<RP, 2>;
    using Smem = TileSmem<T, MEQN, MAUX, R, FLUXFORM, REGS>;
    constexpr int QW = Smem::QW, QS = Smem::QS, QP = Smem::QP, GS = Smem::GS;
    constexpr int L = T + 4;               // lanes per row/column: one per staged cell
    constexpr int SEG = 64 / L;            // rows (columns) per wave
    constexpr int NC = cell_values<RP>::N; // per-cell values of the Riemann solver
    constexpr int NT = NW * 64;
    constexpr int NQ = MEQN * QW * QW;     // q values staged per tile
    constexpr int NA = MAUX * QW * QW;     // aux values staged per tile
    constexpr int KQ = (NQ + NT - 1) / NT; // per-thread prefetch registers
    constexpr int KA = MAUX > 0 ? (NA + NT - 1) / NT : 1;
    static_assert(SEG >= 1, "tile too wide for one wave");
    __shared__ Smem sm;
    R* const s_fm = sm.fm();
    R* const s_fp = sm.fp();
    R* const s_gt = sm.gt();
    R* const s_ab = sm.ab();

    const int ntiles = a.nwork * a.ntx * a.nty;
    const int mx = a.mx, my = a.my, mbc = a.mbc;
    const int nxp = mx + 2 * mbc, nyp = my + 2 * mbc;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // lane (seg, kp): cell kp of a staged row (column) and the interface
    // k = kp - 1 between cells kp - 1 and kp; the left cell's values come from
    // the lane before (lane kp = 0 has no interface of its own)
    const int seg = lane / L, kp = lane - seg * L, k = kp - 1;
    double cfl = 0.0;

    auto tile_slot = [&](int tile, int& tx, int& ty) {
        const int tpp = a.ntx * a.nty;
        const int wid = tile / tpp, tix = tile - wid * tpp;
        tx = tix % a.ntx;
        ty = tix / a.ntx;
        return a.work[wid];
    };
    // Cells before the tile in x (y): tiles step by T, and when T does not
    // divide mx (my) the last tile is moved back to end at the patch edge, so
    // it overlaps its neighbour.  Both compute the shared cells from the same
    // inputs with the same code and write the same bits; registers along a
    // patch edge are written by the tile that owns the position (the first).
    auto origin_x = [&](int tx) { return min(tx * T, mx - T); };
    auto origin_y = [&](int ty) { return min(ty * T, my - T); };
    // Staging map, the same for every tile: element e = threadIdx.x + kk*NT of
    // the (field m, tile row qj, tile column qi) halo tile sits at goff[kk] +
    // (tile origin) in the patch slab and at loff[kk] in LDS.  Every chunk kk
    // but a partial last one holds an element in every thread (the test is
    // then a compile-time constant, and no per-chunk lane mask lives through
    // the tile loop).
    auto has_q = [&](int kk) { return (kk + 1) * NT <= NQ || (int) threadIdx.x + kk * NT < NQ; };
    auto has_a = [&](int kk) { return MAUX > 0 && ((kk + 1) * NT <= NA || (int) threadIdx.x + kk * NT < NA); };
    int goff[KQ], loff[KQ];
#pragma unroll
    for (int kk = 0; kk < KQ; kk++) {
        const int e = threadIdx.x + kk * NT;
        const int m = e / (QW * QW), rem = e - m * QW * QW;
        const int qj = rem / QW, qi = rem - qj * QW;
        // cell (i, j) = (i0 - 2 + qi, j0 - 2 + qj) of field m
        goff[kk] = (m * nyp + qj + mbc - 2) * nxp + qi + mbc - 2;
        loff[kk] = m * QP + qj * QS + qi;
    }
    // Deferred fix-up: entries (face, field m, position 0..QW-1 along the strip)
    // of the tile's four interior edge strips.
    constexpr int NEDGE = 4 * MEQN * QW;
    int gaoff[KA], laoff[KA];
#pragma unroll
    for (int kk = 0; kk < KA; kk++) {
        const int e = threadIdx.x + kk * NT;
        const int m = e / (QW * QW), rem = e - m * QW * QW;
        const int qj = rem / QW, qi = rem - qj * QW;
        gaoff[kk] = (m * nyp + qj + mbc - 2) * nxp + qi + mbc - 2;
        laoff[kk] = m * QP + qj * QS + qi;
    }

    // Halo column and row of each staged element: code = cx + 5 cy, cx = 1, 2
    // for tile columns 0, 1 (left of the tile), 3, 4 for columns T+2, T+3
    // (right of it), 0 inside; cy likewise for rows.  Five bits per element;
    // s_pull[code] is the tile's offset for that column and row: a patch face's
    // or corner's where the cell is a ghost cell of the patch, 0 for the patch's
    // own cells.  Any tile's halo can reach ghost cells, not only the first and
    // last tile's (an overlapped last tile leaves the one before it within two
    // cells of the patch edge).
    constexpr int NPC = (KQ + 11) / 12;   // 12 codes per 64-bit word
    unsigned long long pcode[NPC];
#pragma unroll
    for (int w = 0; w < NPC; w++) pcode[w] = 0;
#pragma unroll
    for (int kk = 0; kk < KQ; kk++) {
        const int e = threadIdx.x + kk * NT;
        const int m = e / (QW * QW), rem = e - m * QW * QW;
        const int qj = rem / QW, qi = rem - qj * QW;
        const int cx = qi < 2 ? 1 + qi : qi > T + 1 ? qi - T + 1 : 0;
        const int cy = qj < 2 ? 1 + qj : qj > T + 1 ? qj - T + 1 : 0;
        if (e < NQ) pcode[kk / 12] |= (unsigned long long) (cx + 5 * cy) << (5 * (kk % 12));
    }
    __shared__ int s_pull[25];

    // prefetch registers for the next tile
    double pq[KQ];
    R pa[KA];
    auto prefetch = [&](int tile) {
        int tx, ty;
        const int slot = tile_slot(tile, tx, ty);
        const long long shift = (long long) origin_y(ty) * nxp + origin_x(tx);  // tile origin in the slab
        const double* __restrict__ qp = a.qsrc + (long long) slot * a.pstride + shift;
        if (a.pull) {
#pragma unroll
            for (int kk = 0; kk < KQ; kk++)
                pq[kk] = has_q(kk) ? qp[goff[kk] + s_pull[(pcode[kk / 12] >> (5 * (kk % 12))) & 31]] : 0.0;
        } else {
#pragma unroll
            for (int kk = 0; kk < KQ; kk++) pq[kk] = has_q(kk) ? qp[goff[kk]] : 0.0;
        }
        if constexpr (MAUX > 0) {
            const double* __restrict__ ap = a.aux + (long long) slot * a.astride + shift;
#pragma unroll
            for (int kk = 0; kk < KA; kk++) pa[kk] = has_a(kk) ? (R) ap[gaoff[kk]] : R(0);
        }
    };

    auto SQ = [&](int m, int qi, int qj) -> R { return (R) sm.q[m * QP + qj * QS + qi]; };
    auto SQD = [&](int m, int qi, int qj) -> double { return sm.q[m * QP + qj * QS + qi]; };
    auto SA = [&](int m, int qi, int qj) -> R {
        if constexpr (MAUX > 0) return sm.aux[m * QP + qj * QS + qi];
        else return R(0);
    };
    // the MAUX aux values of one staged cell (what a solver does not read is
    // never loaded)
    constexpr int MAD = MAUX > 0 ? MAUX : 1;
    auto aux_cell = [&](R* v, int qi, int qj) {
#pragma unroll
        for (int m = 0; m < MAD; m++) v[m] = SA(m, qi, qj);
    };
    // ab[w][m][r][a]
    auto AB = [&](int w, int m, int r, int c) -> R& {
        return s_ab[((w * MEQN + m) * (T + 2) + r) * GS + c];
    };
    // gt[m][t][a]
    auto GT = [&](int m, int t, int c) -> R& { return s_gt[(m * (T + 1) + t) * GS + c]; };

    // Workgroups are dispatched round-robin over the 8 XCDs.  Renumber them so
    // that, in every round of the persistent loop, each XCD works on one
    // contiguous run of tiles: neighbouring tiles of a patch share halo cells,
    // which then come from the same L2.
    // The offsets of tile t for s_pull (threads 0..24, one halo column and
    // row class each): a ghost cell beside a patch face takes that face's
    // offset, one beside a patch corner the corner's, a cell of the patch 0.
    auto pull_table = [&](int t) {
        if (threadIdx.x < 25) {
            int tx, ty;
            const int slot = tile_slot(t, tx, ty);
            const int* __restrict__ pt = a.pull + 8 * (long long) slot;
            const int cx = threadIdx.x % 5, cy = threadIdx.x / 5;
            // the class's patch column and row (1-based; cells 1..mx x 1..my)
            const int i = origin_x(tx) + (cx == 0 ? 1 : cx < 3 ? cx - 2 : T + cx - 2);
            const int j = origin_y(ty) + (cy == 0 ? 1 : cy < 3 ? cy - 2 : T + cy - 2);
            const bool gx = i < 1 || i > mx, gy = j < 1 || j > my;
            const int hx = i > mx ? 1 : 0, hy = j > my ? 1 : 0;
            int d = 0;
            if (gx && gy) d = pt[4 + hx + 2 * hy];
            else if (gx) d = pt[hx];
            else if (gy) d = pt[2 + hy];
            s_pull[threadIdx.x] = d;
        }
    };

    const int G = gridDim.x;
    int tile = (G % 8 == 0) ? (int) (blockIdx.x % 8) * (G / 8) + (int) (blockIdx.x / 8) : (int) blockIdx.x;
    if (a.pull && tile < ntiles) {
        pull_table(tile);
        __syncthreads();
    }
    if (tile < ntiles) prefetch(tile);

    for (; tile < ntiles; tile += G) {
        int tx, ty;
        const int slot = tile_slot(tile, tx, ty);
        const int i0 = 1 + origin_x(tx), j0 = 1 + origin_y(ty);  // first interior cell of the tile (1-based)
        // cells of an overlapping last tile that the tile before it owns
        const int own_x = tx * T + 1 - i0, own_y = ty * T + 1 - j0;
        const double dt = a.dtscale ? a.dt * a.dtscale[slot] : a.dt;
        const double dtdx64 = dt / a.dxs[slot], dtdy64 = dt / a.dys[slot];
        const R dtdx = (R) dtdx64, dtdy = (R) dtdy64;
        const bool first_x = tx == 0, last_x = tx == a.ntx - 1;
        const bool first_y = ty == 0, last_y = ty == a.nty - 1;

        // the next tile's offsets, read by its prefetch after the staging barrier
        if (a.pull && tile + G < ntiles) pull_table(tile + G);
        // ---- stage 0: the prefetched tile + 2-cell halo into LDS ----
#pragma unroll
        for (int kk = 0; kk < KQ; kk++)
            if (has_q(kk)) sm.q[loff[kk]] = pq[kk];
        if constexpr (MAUX > 0) {
#pragma unroll
            for (int kk = 0; kk < KA; kk++)
                if (has_a(kk)) sm.aux[laoff[kk]] = pa[kk];
        }
        __syncthreads();
        if (a.qcorr) {
            // the pending fix-up of the last synchronisation: each corrected edge
            // cell of the staged tile plus its increments, in the reference's
            // order (one thread per cell; a corner cell belongs to its x-face
            // entry when that face is corrected).  A face's edge cells are staged
            // by every tile whose halo reaches them: the first and last tile
            // along it, and the one before an overlapped last tile when the
            // overlap leaves it within two cells of the edge.
            const int mask = a.cmask[slot];
            const int tb = ((int) (i0 <= 3) | ((int) (i0 + T + 1 >= mx) << 1) | ((int) (j0 <= 3) << 2) |
                            ((int) (j0 + T + 1 >= my) << 3)) & mask;
            if (tb) {
                const double* __restrict__ qc = a.qcorr + (long long) slot * a.cstride;
#pragma unroll 1
                for (int x = threadIdx.x; x < NEDGE; x += NT) {
                    const int face = x / (MEQN * QW), r = x - face * (MEQN * QW);
                    const int m = r / QW, q = r - m * QW;   // tile index along the strip
                    if (!((tb >> face) & 1)) continue;
                    // patch cell along the strip: j (faces 0, 1) or i (faces 2, 3), 1-based
                    const int len = face < 2 ? my : mx;
                    const int c = (face < 2 ? j0 : i0) + q - 2;
                    if (c < 1 || c > len) continue;
                    const int other = (c == 1 || c == len) ? (face < 2 ? (c == 1 ? 2 : 3) : (c == 1 ? 0 : 1)) : -1;
                    const bool other_on = other >= 0 && ((mask >> other) & 1);
                    if (face >= 2 && other_on) continue;
                    // the edge cell's tile column (faces 0, 1) or row (2, 3)
                    const int qi = face == 0 ? 3 - i0 : face == 1 ? mx - i0 + 2 : q;
                    const int qj = face == 2 ? 3 - j0 : face == 3 ? my - j0 + 2 : q;
                    double& cell = sm.q[m * QP + qj * QS + qi];
                    const double v = qc[(face * MEQN + m) * len + c - 1];
                    double y = cell;
                    if (other_on) {
                        const double w = qc[(other * MEQN + m) * mx + (face == 0 ? 0 : mx - 1)];
                        if (sync_y_first(mask, face, other)) { y = y + w; y = y + v; }
                        else { y = y + v; y = y + w; }
                    } else {
                        y = y + v;
                    }
                    cell = y;
                }
            }
            __syncthreads();
        }
        if constexpr (REGS && !FLUXFORM) {
            double* __restrict__ rg = a.reg + (long long) slot * a.rstride;
            const double dtdy_el = dt * a.dys[slot], dtdx_el = dt * a.dxs[slot];
            // time-sync registers: conservative flux of the pre-step state at the first
            // interior and first ghost cell of every edge (clawpack46_time_sync.f:58-147)
            for (int e = threadIdx.x; e < 8 * T; e += NT) {
                const int face = e / (2 * T), rem = e - face * (2 * T);
                const int kio = rem / T, c = rem - kio * T;
                bool on;
                int qi, qj;
                if (face == 0) { on = first_x; qi = 2 - kio; qj = c + 2; }
                else if (face == 1) { on = last_x; qi = T + 1 + kio; qj = c + 2; }
                else if (face == 2) { on = first_y; qi = c + 2; qj = 2 - kio; }
                else { on = last_y; qi = c + 2; qj = T + 1 + kio; }
                if (!on || c < (face < 2 ? own_y : own_x)) continue;
                double qcell[MEQN], acell[MAUX > 0 ? MAUX : 1], fl[MEQN];
#pragma unroll
                for (int m = 0; m < MEQN; m++) qcell[m] = SQD(m, qi, qj);
#pragma unroll
                for (int m = 0; m < (MAUX > 0 ? MAUX : 1); m++) acell[m] = (double) SA(m, qi, qj);
                if (face < 2) RP::template consflux<1>(qcell, acell, a.pp, fl);
                else RP::template consflux<2>(qcell, acell, a.pp, fl);
                const double el = face < 2 ? dtdy_el : dtdx_el;
                const int pos = (face < 2 ? j0 : i0) - 1 + c;
#pragma unroll
                for (int m = 0; m < MEQN; m++)
                    reg_add(rg + ((face * 4 + REG_EFIN + kio) * MEQN + m) * mx + pos, el * fl[m], a.reg_accumulate);
            }
        }
        // the next tile's data travels while this one is computed
        if (tile + G < ntiles) prefetch(tile + G);

        // ============================ x sweep ============================
        // lane (seg, k): row j = j0-1+r, interface i = i0-1+k between cells i-1 and i.
        // The Courant number max |dtdx s| is dtdx * max |s| exactly (rounding is
        // monotone), so a lane keeps max |s| and scales it once per sweep.
        R smax = R(0);
        for (int rbase = wave * SEG; rbase < T + 2; rbase += NW * SEG) {
            const int r = rbase + seg;
            const bool act = (seg < SEG) && (r < T + 2);
            const int rc = r < T + 1 ? r : T + 1;
            const int kc = kp < L ? kp : L - 1;     // the lane's cell (column)
            const int kl = kc > 0 ? kc - 1 : 0;     // the cell left of its interface
            R qL[MEQN], qR[MEQN], auxL[MAD], auxR[MAD];
#pragma unroll
            for (int m = 0; m < MEQN; m++) qR[m] = SQ(m, kc, rc + 1);
#pragma unroll
            for (int m = 0; m < MEQN; m++) qL[m] = prev_lane(qR[m]);
            aux_cell(auxL, kl, rc + 1);
            aux_cell(auxR, kc, rc + 1);
            R wave_[MW][MEQN], s[MW], amdq[MEQN], apdq[MEQN], roe[NROE];
            riemann<RP, 1, NC>(qL, qR, auxL, auxR, a.pp, wave_, s, amdq, apdq, roe);
            // The reference takes the CFL over x-interfaces i = 1..mx+1 of rows
            // j = 2-mbc..my+mbc-1 (clawpack46_flux2.f:129-137, clawpack46_step2.f:113).
            // A tile's interfaces k = 1..T+1 of its rows r = 0..T+1 lie in that set
            // and the tiles cover it; an interface two tiles share has the same
            // speeds in both, so the maximum is the reference's.
            if (act && k >= 1 && k <= T + 1) {
#pragma unroll
                for (int mw = 0; mw < MW; mw++) smax = fmax(smax, fabs(s[mw]));
            }
            R cqxx[MEQN];
#pragma unroll
            for (int m = 0; m < MEQN; m++) cqxx[m] = R(0);
            if (a.order2) {
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    const int mth = LIM > 0 ? LIM : a.mthlim[mw];
                    if (mth == 0) continue;
                    R wn = R(0), dotl = R(0);
#pragma unroll
                    for (int m = 0; m < MEQN; m++)
                        if (!WZ1::at(mw, m)) wn = wn + wave_[mw][m] * wave_[mw][m];
#pragma unroll
                    for (int m = 0; m < MEQN; m++) {
                        if (WZ1::at(mw, m)) continue;
                        const R wl = prev_lane(wave_[mw][m]);
                        dotl = dotl + wl * wave_[mw][m];
                    }
                    const R dotr = next_lane(dotl);
                    // phi = 1 where the reference leaves the wave as it is (zero
                    // norm, or an interface whose flux no cell of the tile reads):
                    // phi * w = w bit for bit
                    const R rr = dv(s[mw] > R(0) ? dotl : dotr, wn);
                    const R phi = (k >= 1 && k <= T + 1 && wave_nonzero(wn)) ? philim<R>(mth, rr) : R(1);
#pragma unroll
                    for (int m = 0; m < MEQN; m++)
                        if (!WZ1::at(mw, m)) wave_[mw][m] = phi * wave_[mw][m];
                }
#pragma unroll
                for (int m = 0; m < MEQN; m++) {
                    R c = R(0);
#pragma unroll
                    for (int mw = 0; mw < MW; mw++)
                        if (!WZ1::at(mw, m)) c = c + fabs(s[mw]) * (R(1) - fabs(s[mw]) * dtdx) * wave_[mw][m];
                    cqxx[m] = c;
                }
            }
            if (act && r >= 1 && r <= T && k >= 1 && k <= T + 1) {
#pragma unroll
                for (int m = 0; m < MEQN; m++) {
                    s_fm[(m * T + (r - 1)) * (T + 1) + (k - 1)] = zadd(amdq[m]) + R(0.5) * cqxx[m];
                    if constexpr (!FLUXFORM)
                        s_fp[(m * T + (r - 1)) * (T + 1) + (k - 1)] = zsub(apdq[m]) + R(0.5) * cqxx[m];
                }
            }
            R tx1[MEQN], tx2[MEQN];
            if (a.transverse) {
                if (a.order2 && a.transverse == 2) {
#pragma unroll
                    for (int m = 0; m < MEQN; m++) { amdq[m] = amdq[m] + cqxx[m]; apdq[m] = apdq[m] - cqxx[m]; }
                }
                // FP32 with a linear rpt2: the factor dt / (2 dx) on asdq
                constexpr bool PRE = sizeof(R) == 4 && rpt2_linear<RP>::value;
                const R half = R(0.5) * dtdx;
                if constexpr (PRE) {
#pragma unroll
                    for (int m = 0; m < MEQN; m++) { amdq[m] = half * amdq[m]; apdq[m] = half * apdq[m]; }
                }
                R bmA[MEQN], bpA[MEQN], bmP[MEQN], bpP[MEQN];
                // aux1, aux2, aux3: rows j-1, j, j+1 of column i1 = i-1 (amdq) or i (apdq)
                {
                    R a1[MAD], a2[MAD], a3[MAD];
                    aux_cell(a1, kl, rc); aux_cell(a2, kl, rc + 1); aux_cell(a3, kl, rc + 2);
                    RP::template rpt2<1>(roe, amdq, a1, a2, a3, a.pp, bmA, bpA);
                    aux_cell(a1, kc, rc); aux_cell(a2, kc, rc + 1); aux_cell(a3, kc, rc + 2);
                    RP::template rpt2<1>(roe, apdq, a1, a2, a3, a.pp, bmP, bpP);
                }
#pragma unroll
                for (int m = 0; m < MEQN; m++) {
                    const R gAd = PRE ? bmA[m] : half * bmA[m], gAu = PRE ? bpA[m] : half * bpA[m];
                    const R gPd = PRE ? bmP[m] : half * bmP[m], gPu = PRE ? bpP[m] : half * bpP[m];
                    tx1[m] = zsub(next_lane(gAd)) - gPd;
                    tx2[m] = zsub(next_lane(gAu)) - gPu;
                }
            } else {
#pragma unroll
                for (int m = 0; m < MEQN; m++) { tx1[m] = R(0); tx2[m] = R(0); }
            }
            if (act && k >= 1 && k <= T) {
#pragma unroll
                for (int m = 0; m < MEQN; m++) { AB(0, m, r, k - 1) = tx1[m]; AB(1, m, r, k - 1) = tx2[m]; }
            }
        }
        __syncthreads();

        // gm(i,j) = gp(i,j) after the x sweeps: row j-1 adds gadd(.,2), then row j adds gadd(.,1)
        if constexpr (NT % T == 0) {
            // thread (t0, c): column c, rows t0, t0 + NT/T, ... (no index division)
            constexpr int TS = NT / T;
            const int c = threadIdx.x % T, t0 = threadIdx.x / T;
#pragma unroll
            for (int m = 0; m < MEQN; m++)
                for (int t = t0; t <= T; t += TS) GT(m, t, c) = zadd(AB(1, m, t, c)) + AB(0, m, t + 1, c);
        } else {
            for (int e = threadIdx.x; e < MEQN * (T + 1) * T; e += NT) {
                const int m = e / ((T + 1) * T), rem = e - m * (T + 1) * T;
                const int t = rem / T, c = rem - t * T;
                GT(m, t, c) = zadd(AB(1, m, t, c)) + AB(0, m, t + 1, c);
            }
        }
        __syncthreads();
        cfl = fmax(cfl, dtdx64 * (double) smax);

        // ============================ y sweep ============================
        // lane (seg, k): column i = i0-1+c, interface j = j0-1+k between cells j-1 and j
        smax = R(0);
        for (int cbase = wave * SEG; cbase < T + 2; cbase += NW * SEG) {
            const int c = cbase + seg;
            const bool act = (seg < SEG) && (c < T + 2);
            const int cc = c < T + 1 ? c : T + 1;
            const int kc = kp < L ? kp : L - 1;     // the lane's cell (row)
            const int kl = kc > 0 ? kc - 1 : 0;     // the cell below its interface
            R qL[MEQN], qR[MEQN], auxL[MAD], auxR[MAD];
#pragma unroll
            for (int m = 0; m < MEQN; m++) qR[m] = SQ(m, cc + 1, kc);
#pragma unroll
            for (int m = 0; m < MEQN; m++) qL[m] = prev_lane(qR[m]);
            aux_cell(auxL, cc + 1, kl);
            aux_cell(auxR, cc + 1, kc);
            R wave_[MW][MEQN], s[MW], amdq[MEQN], apdq[MEQN], roe[NROE];
            riemann<RP, 2, NC>(qL, qR, auxL, auxR, a.pp, wave_, s, amdq, apdq, roe);
            if (act && k >= 1 && k <= T + 1) {   // as in the x sweep
#pragma unroll
                for (int mw = 0; mw < MW; mw++) smax = fmax(smax, fabs(s[mw]));
            }
            R cqxx[MEQN];
#pragma unroll
            for (int m = 0; m < MEQN; m++) cqxx[m] = R(0);
            if (a.order2) {
#pragma unroll
                for (int mw = 0; mw < MW; mw++) {
                    const int mth = LIM > 0 ? LIM : a.mthlim[mw];
                    if (mth == 0) continue;
                    R wn = R(0), dotl = R(0);
#pragma unroll
                    for (int m = 0; m < MEQN; m++)
                        if (!WZ2::at(mw, m)) wn = wn + wave_[mw][m] * wave_[mw][m];
#pragma unroll
                    for (int m = 0; m < MEQN; m++) {
                        if (WZ2::at(mw, m)) continue;
                        const R wl = prev_lane(wave_[mw][m]);
                        dotl = dotl + wl * wave_[mw][m];
                    }
                    const R dotr = next_lane(dotl);
                    // phi = 1 where the reference leaves the wave as it is (zero
                    // norm, or an interface whose flux no cell of the tile reads):
                    // phi * w = w bit for bit
                    const R rr = dv(s[mw] > R(0) ? dotl : dotr, wn);
                    const R phi = (k >= 1 && k <= T + 1 && wave_nonzero(wn)) ? philim<R>(mth, rr) : R(1);
#pragma unroll
                    for (int m = 0; m < MEQN; m++)
                        if (!WZ2::at(mw, m)) wave_[mw][m] = phi * wave_[mw][m];
                }
#pragma unroll
                for (int m = 0; m < MEQN; m++) {
                    R cq = R(0);
#pragma unroll
                    for (int mw = 0; mw < MW; mw++)
                        if (!WZ2::at(mw, m)) cq = cq + fabs(s[mw]) * (R(1) - fabs(s[mw]) * dtdy) * wave_[mw][m];
                    cqxx[m] = cq;
                }
            }
            // normal y-fluxes: gm = gt + faddm, gp = gt + faddp
            const bool own_col = c >= 1 && c <= T;
            if constexpr (FLUXFORM) {
                if (act && own_col && k >= 1 && k <= T + 1) {
#pragma unroll
                    for (int m = 0; m < MEQN; m++) {
                        R& g = GT(m, k - 1, c - 1);
                        g = g + (zadd(amdq[m]) + R(0.5) * cqxx[m]);
                    }
                }
            } else {
                R gm[MEQN], gp[MEQN];
                const int kk = (k >= 1 && k <= T + 1) ? k - 1 : 0;
                const int ccol = own_col ? c - 1 : 0;
#pragma unroll
                for (int m = 0; m < MEQN; m++) {
                    const R g = GT(m, kk, ccol);
                    gm[m] = g + (zadd(amdq[m]) + R(0.5) * cqxx[m]);
                    gp[m] = g + (zsub(apdq[m]) + R(0.5) * cqxx[m]);
                }
#pragma unroll
                for (int m = 0; m < MEQN; m++) {
                    const R gm_up = next_lane(gm[m]);
                    if (act && own_col && k >= 1 && k <= T)
                        GT(m, k - 1, c - 1) = gm_up - gp[m];  // Dy(i, j)
                }
                // gm, gp at the tile's bottom (t = 0) and top (t = T) y-interfaces,
                // kept in LDS for the time-sync registers of the patch edges
                if constexpr (REGS) {
                    if (act && own_col && (k == 1 || k == T + 1)) {
                        const int side = k == 1 ? 0 : 1;
#pragma unroll
                        for (int m = 0; m < MEQN; m++) {
                            sm.eg[((side * 2 + 0) * MEQN + m) * T + c - 1] = gm[m];
                            sm.eg[((side * 2 + 1) * MEQN + m) * T + c - 1] = gp[m];
                        }
                    }
                }
            }
            R ty1[MEQN], ty2[MEQN];
            if (a.transverse) {
                if (a.order2 && a.transverse == 2) {
#pragma unroll
                    for (int m = 0; m < MEQN; m++) { amdq[m] = amdq[m] + cqxx[m]; apdq[m] = apdq[m] - cqxx[m]; }
                }
                // FP32 with a linear rpt2: the factor dt / (2 dx) on asdq
                constexpr bool PRE = sizeof(R) == 4 && rpt2_linear<RP>::value;
                const R half = R(0.5) * dtdy;
                if constexpr (PRE) {
#pragma unroll
                    for (int m = 0; m < MEQN; m++) { amdq[m] = half * amdq[m]; apdq[m] = half * apdq[m]; }
                }
                R bmA[MEQN], bpA[MEQN], bmP[MEQN], bpP[MEQN];
                // aux1, aux2, aux3: columns i-1, i, i+1 of row j1 = j-1 (amdq) or j (apdq)
                {
                    R a1[MAD], a2[MAD], a3[MAD];
                    aux_cell(a1, cc, kl); aux_cell(a2, cc + 1, kl); aux_cell(a3, cc + 2, kl);
                    RP::template rpt2<2>(roe, amdq, a1, a2, a3, a.pp, bmA, bpA);
                    aux_cell(a1, cc, kc); aux_cell(a2, cc + 1, kc); aux_cell(a3, cc + 2, kc);
                    RP::template rpt2<2>(roe, apdq, a1, a2, a3, a.pp, bmP, bpP);
                }
#pragma unroll
                for (int m = 0; m < MEQN; m++) {
                    const R gAd = PRE ? bmA[m] : half * bmA[m], gAu = PRE ? bpA[m] : half * bpA[m];
                    const R gPd = PRE ? bmP[m] : half * bmP[m], gPu = PRE ? bpP[m] : half * bpP[m];
                    ty1[m] = zsub(next_lane(gAd)) - gPd;
                    ty2[m] = zsub(next_lane(gAu)) - gPu;
                }
            } else {
#pragma unroll
                for (int m = 0; m < MEQN; m++) { ty1[m] = R(0); ty2[m] = R(0); }
            }
            if (act && k >= 1 && k <= T) {
#pragma unroll
                for (int m = 0; m < MEQN; m++) { AB(0, m, c, k - 1) = ty1[m]; AB(1, m, c, k - 1) = ty2[m]; }
            }
        }
        __syncthreads();
        cfl = fmax(cfl, dtdy64 * (double) smax);

        double* __restrict__ qo = a.qdst + (long long) slot * a.pstride;
        // fm(i,j) = ((0 + faddm_x(i,j)) + ty2(i-1,j)) + ty1(i,j), same for fp
        if constexpr (!FLUXFORM) {
            for (int e = threadIdx.x; e < MEQN * T * T; e += NT) {
                const int m = e / (T * T), rem = e - m * T * T;
                const int b = rem / T, c = rem - b * T;   // cell i = i0+c, j = j0+b
                const R fmR = (zadd(s_fm[(m * T + b) * (T + 1) + c + 1]) + AB(1, m, c + 1, b)) + AB(0, m, c + 2, b);
                const R fpL = (zadd(s_fp[(m * T + b) * (T + 1) + c]) + AB(1, m, c, b)) + AB(0, m, c + 1, b);
                const R Dx = fmR - fpL;
                const R Dy = GT(m, b, c);
                const double qc = SQD(m, c + 2, b + 2);
                const int i = i0 + c, j = j0 + b;
                const long long off = ((long long) m * nyp + (j + mbc - 1)) * nxp + (i + mbc - 1);
                if constexpr (sizeof(R) == sizeof(double))
                    qo[off] = (qc - dtdx * Dx) - dtdy * Dy;
                else  // FP32 fluctuations, FP64 state: q never passes through FP32
                    qo[off] = (qc - dtdx64 * (double) Dx) - dtdy64 * (double) Dy;
            }
            if constexpr (REGS) {
                double* __restrict__ rg = a.reg + (long long) slot * a.rstride;
                const double dtdy_el = dt * a.dys[slot], dtdx_el = dt * a.dxs[slot];
                // fm, fp at the patch's left (i = 1) and right (i = mx+1) edges,
                // summed as the update sums them (clawpack46_time_sync.f:192-199)
                for (int e = threadIdx.x; e < 2 * MEQN * T; e += NT) {
                    const int side = e / (MEQN * T), rem = e - side * (MEQN * T);
                    const int m = rem / T, b = rem - m * T;
                    if ((side == 0 ? !first_x : !last_x) || b < own_y) continue;
                    const int s0 = side ? T : 0;
                    const R fmv = (zadd(s_fm[(m * T + b) * (T + 1) + s0]) + AB(1, m, s0, b)) + AB(0, m, s0 + 1, b);
                    const R fpv = (zadd(s_fp[(m * T + b) * (T + 1) + s0]) + AB(1, m, s0, b)) + AB(0, m, s0 + 1, b);
                    reg_add(rg + ((side * 4 + REG_FM) * MEQN + m) * mx + j0 - 1 + b, dtdy_el * (double) fmv, a.reg_accumulate);
                    reg_sub(rg + ((side * 4 + REG_FP) * MEQN + m) * mx + j0 - 1 + b, dtdy_el * (double) fpv, a.reg_accumulate);
                }
                // gm, gp at the patch's bottom (j = 1) and top (j = my+1) edges
                // (clawpack46_time_sync.f:201-209)
                for (int e = threadIdx.x; e < 2 * MEQN * T; e += NT) {
                    const int side = e / (MEQN * T), rem = e - side * (MEQN * T);
                    const int m = rem / T, c = rem - m * T;
                    if ((side == 0 ? !first_y : !last_y) || c < own_x) continue;
                    const int face = 2 + side, pos = i0 - 1 + c;
                    reg_add(rg + ((face * 4 + REG_FM) * MEQN + m) * mx + pos,
                            dtdx_el * (double) sm.eg[((side * 2 + 0) * MEQN + m) * T + c], a.reg_accumulate);
                    reg_sub(rg + ((face * 4 + REG_FP) * MEQN + m) * mx + pos,
                            dtdx_el * (double) sm.eg[((side * 2 + 1) * MEQN + m) * T + c], a.reg_accumulate);
                }
            }
        } else {
            // conservative FP64 interface fluxes F = f(q(i-1,j)) + fm(i,j) and
            // G = g(q(i,j-1)) + gm(i,j).  Each cell evaluates the four it needs
            // (RP::fluxes is compiled without contraction), so the two cells on
            // either side of an interface compute the same bits from the same
            // LDS values and the FP64 update telescopes exactly; no flux array is
            // kept in LDS.
            auto xflux = [&](int sx, int b, int m, double f) -> double {  // x-interface sx at row b
                const R fmv = (zadd(s_fm[(m * T + b) * (T + 1) + sx]) + AB(1, m, sx, b)) + AB(0, m, sx + 1, b);
                return f + (double) fmv;
            };
            auto yflux = [&](int t, int c, int m, double g) -> double {   // y-interface t at column c
                return g + (double) GT(m, t, c);
            };
            auto cell_fluxes = [&](int qi, int qj, double* f, double* g) {
                double qv[MEQN], av[MAUX > 0 ? MAUX : 1];
#pragma unroll
                for (int m = 0; m < MEQN; m++) qv[m] = SQD(m, qi, qj);
#pragma unroll
                for (int m = 0; m < (MAUX > 0 ? MAUX : 1); m++) av[m] = (double) SA(m, qi, qj);
                RP::fluxes(qv, av, a.pp, f, g);
            };
            if constexpr (REGS) {
                // the shared FP64 flux through each patch edge, times dt * edge length
                double* __restrict__ rg = a.reg + (long long) slot * a.rstride;
                const double dtdy_el = dt * a.dys[slot], dtdx_el = dt * a.dxs[slot];
                for (int e = threadIdx.x; e < 4 * T; e += NT) {
                    const int face = e / T, c = e - face * T;
                    if (face == 0 ? !first_x : face == 1 ? !last_x : face == 2 ? !first_y : !last_y) continue;
                    if (c < (face < 2 ? own_y : own_x)) continue;
                    double f[MEQN], g[MEQN];
                    if (face < 2) cell_fluxes(face == 0 ? 1 : T + 1, c + 2, f, g);
                    else cell_fluxes(c + 2, face == 2 ? 1 : T + 1, f, g);
                    const int pos = (face < 2 ? j0 : i0) - 1 + c;
#pragma unroll
                    for (int m = 0; m < MEQN; m++) {
                        const double F = face == 0 ? xflux(0, c, m, f[m]) : face == 1 ? xflux(T, c, m, f[m])
                                       : face == 2 ? yflux(0, c, m, g[m]) : yflux(T, c, m, g[m]);
                        reg_add(rg + ((face * 4 + REG_EFIN) * MEQN + m) * mx + pos,
                                (face < 2 ? dtdy_el : dtdx_el) * F, a.reg_accumulate);
                    }
                }
            }
            for (int e = threadIdx.x; e < T * T; e += NT) {
                const int b = e / T, c = e - b * T;   // cell i = i0+c, j = j0+b
                double fc[MEQN], gc[MEQN], fl[MEQN], gl[MEQN], fb[MEQN], gb[MEQN];
                cell_fluxes(c + 2, b + 2, fc, gc);    // this cell
                cell_fluxes(c + 1, b + 2, fl, gl);    // its left neighbour
                cell_fluxes(c + 2, b + 1, fb, gb);    // the one below
                const int i = i0 + c, j = j0 + b;
#pragma unroll
                for (int m = 0; m < MEQN; m++) {
                    const long long off = ((long long) m * nyp + (j + mbc - 1)) * nxp + (i + mbc - 1);
                    const double qc = SQD(m, c + 2, b + 2);
                    const double dfx = xflux(c + 1, b, m, fc[m]) - xflux(c, b, m, fl[m]);
                    const double dgy = yflux(b + 1, c, m, gc[m]) - yflux(b, c, m, gb[m]);
                    qo[off] = (qc - dtdx64 * dfx) - dtdy64 * dgy;
                }
            }
        }
        __syncthreads();  // LDS is rewritten by the next tile
    }

    // ---- CFL: wave64 max over all tiles of this workgroup, one atomic per wave ----
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) cfl = fmax(cfl, __shfl_xor(cfl, off, 64));
    if (lane == 0 && cfl > 0.0) atomicMax(a.cfl, (unsigned long long) __double_as_longlong(cfl));
}

}  // namespace hipclaw
