// Device Riemann solvers for the batched patch update.
//
// Each solver works on ONE interface held in registers: qL is the cell to the
// left of the interface (the reference's qr(i-1,:)), qR the cell to the right
// (ql(i,:)).  IXY = 1 for an x-interface, 2 for a y-interface.  The arithmetic
// is written in the same order as the reference Fortran so that, compiled
// with FMA contraction off, the FP64 instantiation reproduces it bit for bit:
//
//   shallow water : applications/clawpack/shallow/2d/rp/clawpack46_rpn2.f:91-242,
//                   clawpack46_rpt2.f:287-319 (Roe, entropy fix on)
//   Euler         : applications/clawpack/euler/2d/rp/clawpack46_rpn2_euler4.f:83-285,
//                   clawpack46_rpt2_euler4.f:345-383 (Roe, entropy fix on)
//   advection     : applications/clawpack/advection/2d/all/clawpack46_rpn2adv.f:28-58,
//                   clawpack46_rpt2adv.f:79-86
//
// Solvers are compile-time functors (no device function pointers) so the
// compiler can keep waves, speeds and fluctuations in VGPRs.
#pragma once

#include <type_traits>

#include <hip/hip_runtime.h>

namespace hipclaw {

struct PhysParams {
    double grav;   // shallow water gravity
    double gamma;  // Euler ratio of specific heats
    // A registered solver's own constants (what the reference keeps in a
    // solver's common block or __constant__ memory, e.g. rho, bulk, cc, zz of
    // the acoustics solver's common /cparam/).
    double p[8];
};

// min / max.  FP64 selects as the reference's MIN / MAX (a compare and a
// select); FP32 uses the hardware v_min_f32 / v_max_f32, which differ only
// for NaN operands.
template <typename R>
__device__ __forceinline__ R rmin(R a, R b) { return a < b ? a : b; }
template <typename R>
__device__ __forceinline__ R rmax(R a, R b) { return a > b ? a : b; }
template <>
__device__ __forceinline__ float rmin<float>(float a, float b) { return __builtin_fminf(a, b); }
template <>
__device__ __forceinline__ float rmax<float>(float a, float b) { return __builtin_fmaxf(a, b); }

// Sums that start at zero: 0 + x and 0 - x.  FP64 keeps the reference's
// accumulations from zero, whose only effect is the sign of a zero result;
// FP32 drops them.
template <typename R>
__device__ __forceinline__ R zadd(R x) { return sizeof(R) == 4 ? x : R(0) + x; }
template <typename R>
__device__ __forceinline__ R zsub(R x) { return sizeof(R) == 4 ? -x : R(0) - x; }

// Division and square root.  FP64 keeps IEEE-correct operations (the parity
// form); the FP32-bulk form uses the hardware reciprocal and square root
// (v_rcp_f32, v_sqrt_f32, 1 ulp), whose error is far below the form's
// stated tolerance.
template <typename R>
__device__ __forceinline__ R dv(R a, R b) { return a / b; }

// FP64 reciprocal of the flux form's FP64 fluxes: the hardware estimate
// (v_rcp_f64) refined by two Newton-Raphson steps r <- r + r (1 - x r), which
// square the relative error each.  Not correctly rounded (within a few ulp),
// which the flux form does not need: both cells beside an interface evaluate
// its flux with this same code on the same values, so they get the same bits.
__device__ __forceinline__ double flux_rcp(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    r = __builtin_fma(r, __builtin_fma(-x, r, 1.0), r);
    return r;
}
template <>
__device__ __forceinline__ float dv<float>(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
template <typename R>
__device__ __forceinline__ R sq(R a) { return sqrt(a); }
template <>
__device__ __forceinline__ float sq<float>(float a) { return __builtin_amdgcn_sqrtf(a); }

// A solver whose rpt2 is linear in asdq (RP::RPT2_LINEAR): the FP32 form then
// scales asdq by dt / (2 dx) before the transverse solve instead of scaling
// both results after it.  Solvers without the flag keep the reference's order.
template <class RP, class = void>
struct rpt2_linear : std::false_type {};
template <class RP>
struct rpt2_linear<RP, std::enable_if_t<RP::RPT2_LINEAR>> : std::true_type {};

// ----------------------------------------------------------------- shallow water
// Wave components that are zero by construction.  A sum over the waves that
// starts at +0 may skip those terms without changing its bits (in
// round-to-nearest such a sum is never -0, so adding x * 0 = ±0 leaves it
// as it is), so the step kernel and the solvers skip them.
template <class RP, int IXY, class = void>
struct wave_zeros {
    __host__ __device__ static constexpr bool at(int, int) { return false; }
};
template <class RP, int IXY>
struct wave_zeros<RP, IXY, decltype((void) RP::template wave_zero<IXY>(0, 0))> {
    __host__ __device__ static constexpr bool at(int mw, int m) { return RP::template wave_zero<IXY>(mw, m); }
};

template <typename R>
struct RPShallow {
    static constexpr int MEQN = 3, MWAVES = 3, MAUX = 0;
    static constexpr bool CONSERVATIVE = true;  // amdq + apdq = f(qR) - f(qL)
    static constexpr bool RPT2_LINEAR = true;   // bm, bp linear in asdq
    // u, v, a (and in FP32 1 / (2 a), which rpt2 then takes from the normal solve)
    static constexpr int NROE = sizeof(R) == 4 ? 4 : 3;

    // the shear wave carries no depth and no normal momentum
    template <int IXY>
    __host__ __device__ static constexpr bool wave_zero(int mw, int m) {
        return mw == 1 && (m == 0 || m == (IXY == 1 ? 1 : 2));
    }

    // What rpn2 computes from one cell alone: sqrt(h), hu_n / sqrt(h),
    // hu_t / sqrt(h), u_n = hu_n / h and sqrt(g h) (n: the sweep's normal
    // direction).  The step kernel evaluates them once per cell and hands each
    // interface its two cells' values (the same operations, so the same bits).
    static constexpr int NCELL = 5;
    //
    // FP32 takes all five from one hardware reciprocal square root (v_rsq_f32,
    // 1 ulp): sqrt(h) = h / sqrt(h), 1 / h = (1 / sqrt(h))^2, sqrt(g h) =
    // sqrt(g) sqrt(h); FP64 keeps the reference's divisions and square roots.
    template <int IXY>
    __device__ __forceinline__ static void cell(const R* q, const PhysParams& pp, R c[NCELL]) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        if constexpr (sizeof(R) == 4) {
            const float rs = __builtin_amdgcn_rsqf(q[0]);
            const float hs = q[0] * rs;
            c[0] = hs;
            c[1] = q[mu] * rs;
            c[2] = q[mv] * rs;
            c[3] = c[1] * rs;
            c[4] = sqrt_grav(pp) * hs;
        } else {
            const R hs = sq(q[0]);
            c[0] = hs;
            c[1] = dv(q[mu], hs);
            c[2] = dv(q[mv], hs);
            c[3] = dv(q[mu], q[0]);
            c[4] = sq((R) pp.grav * q[0]);
        }
    }

    // sqrt(g) in FP32 (uniform; the compiler keeps it out of the sweeps' loops)
    __device__ __forceinline__ static float sqrt_grav(const PhysParams& pp) {
        return __builtin_amdgcn_sqrtf((float) pp.grav);
    }

    // u_n - sqrt(g h) (SIGN = -1) or u_n + sqrt(g h) (SIGN = +1) of the
    // intermediate state h, hu_n of the entropy fix
    template <int SIGN>
    __device__ __forceinline__ static R char_speed(R hu, R h, const PhysParams& pp) {
        if constexpr (sizeof(R) == 4) {
            const float rs = __builtin_amdgcn_rsqf(h);
            const float hs = h * rs;
            return SIGN < 0 ? hu * rs * rs - sqrt_grav(pp) * hs : hu * rs * rs + sqrt_grav(pp) * hs;
        } else {
            return SIGN < 0 ? dv(hu, h) - sq((R) pp.grav * h) : dv(hu, h) + sq((R) pp.grav * h);
        }
    }

    template <int IXY>
    __device__ __forceinline__ static void rpn2(const R* qL, const R* qR, const R* auxL,
                                                const R* auxR, const PhysParams& pp,
                                                R wave[MWAVES][MEQN], R s[MWAVES], R amdq[MEQN],
                                                R apdq[MEQN], R roe[NROE]) {
        R cL[NCELL], cR[NCELL];
        cell<IXY>(qL, pp, cL);
        cell<IXY>(qR, pp, cR);
        rpn2c<IXY>(qL, qR, cL, cR, auxL, auxR, pp, wave, s, amdq, apdq, roe);
    }

    // rpn2 from the two cells' values of `cell`
    template <int IXY>
    __device__ __forceinline__ static void rpn2c(const R* qL, const R* qR, const R* cL,
                                                 const R* cR, const R* /*auxL*/,
                                                 const R* /*auxR*/, const PhysParams& pp,
                                                 R wave[MWAVES][MEQN], R s[MWAVES], R amdq[MEQN],
                                                 R apdq[MEQN], R roe[NROE]) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const R grav = (R) pp.grav;
        const R h = dv(qL[0] + qR[0], R(2));
        const R hsqrtl = cL[0];
        const R hsqrtr = cR[0];
        const R hsq2 = hsqrtl + hsqrtr;
        R u, v, a, a1, a3;
        const R d0 = qR[0] - qL[0], d1 = qR[mu] - qL[mu], d2 = qR[mv] - qL[mv];
        if constexpr (sizeof(R) == 4) {
            // one reciprocal and one reciprocal square root: 1 / (2 a) = rsq(g h) / 2
            const float rh = __builtin_amdgcn_rcpf(hsq2);
            u = (cL[1] + cR[1]) * rh;
            v = (cL[2] + cR[2]) * rh;
            const float gh = grav * h, ra = __builtin_amdgcn_rsqf(gh);
            a = gh * ra;
            const float inv2a = 0.5f * ra;
            roe[3] = inv2a;
            a1 = ((u + a) * d0 - d1) * inv2a;
            a3 = (-((u - a) * d0) + d1) * inv2a;
        } else {
            u = dv(cL[1] + cR[1], hsq2);
            v = dv(cL[2] + cR[2], hsq2);
            a = sq(grav * h);
            a1 = dv((u + a) * d0 - d1, R(2) * a);
            a3 = dv(-((u - a) * d0) + d1, R(2) * a);
        }
        roe[0] = u; roe[1] = v; roe[2] = a;
        const R a2 = -(v * d0) + d2;

        wave[0][0] = a1;   wave[0][mu] = a1 * (u - a); wave[0][mv] = a1 * v; s[0] = u - a;
        wave[1][0] = R(0); wave[1][mu] = R(0);         wave[1][mv] = a2;     s[1] = u;
        wave[2][0] = a3;   wave[2][mu] = a3 * (u + a); wave[2][mv] = a3 * v; s[2] = u + a;

        // entropy fix
        const R s0 = cL[3] - cL[4];   // dv(qL[mu], qL[0]) - sq(grav * qL[0])
        if (s0 > R(0) && s[0] > R(0)) {
#pragma unroll
            for (int m = 0; m < 3; m++) amdq[m] = R(0);
        } else {
            const R h1 = qL[0] + wave[0][0];
            const R hu1 = qL[mu] + wave[0][mu];
            const R s1 = char_speed<-1>(hu1, h1, pp);
            R sfract;
            if (s0 < R(0) && s1 > R(0)) sfract = s0 * dv(s1 - s[0], s1 - s0);
            else if (s[0] < R(0)) sfract = s[0];
            else sfract = R(0);
#pragma unroll
            for (int m = 0; m < 3; m++) amdq[m] = sfract * wave[0][m];
            if (!(s[1] > R(0))) {
#pragma unroll
                for (int m = 0; m < 3; m++) amdq[m] = amdq[m] + s[1] * wave[1][m];
                const R s03 = cR[3] + cR[4];   // dv(qR[mu], qR[0]) + sq(grav * qR[0])
                const R h3 = qR[0] - wave[2][0];
                const R hu3 = qR[mu] - wave[2][mu];
                const R s3 = char_speed<1>(hu3, h3, pp);
                bool add = true;
                if (s3 < R(0) && s03 > R(0)) sfract = s3 * dv(s03 - s[2], s03 - s3);
                else if (s[2] < R(0)) sfract = s[2];
                else add = false;
                if (add) {
#pragma unroll
                    for (int m = 0; m < 3; m++) amdq[m] = amdq[m] + sfract * wave[2][m];
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 3; m++) {
            R df = R(0);
#pragma unroll
            for (int mw = 0; mw < 3; mw++)
                if (!wave_zero<IXY>(mw, m)) df = df + s[mw] * wave[mw][m];
            apdq[m] = df - amdq[m];
        }
    }

    template <int IXY>
    __device__ __forceinline__ static void rpt2(const R roe[NROE], const R asdq[MEQN],
                                                const R* /*aux1*/, const R* /*aux2*/,
                                                const R* /*aux3*/, const PhysParams&, R bm[MEQN],
                                                R bp[MEQN]) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const R u = roe[0], v = roe[1], a = roe[2];
        R a1, a3;
        if constexpr (sizeof(R) == 4) {
            a1 = ((v + a) * asdq[0] - asdq[mv]) * roe[3];
            a3 = (-((v - a) * asdq[0]) + asdq[mv]) * roe[3];
        } else {
            a1 = dv((v + a) * asdq[0] - asdq[mv], R(2) * a);
            a3 = dv(-((v - a) * asdq[0]) + asdq[mv], R(2) * a);
        }
        const R a2 = asdq[mu] - u * asdq[0];
        R wb[3][3], sb[3];
        wb[0][0] = a1;   wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a); sb[0] = v - a;
        wb[1][0] = R(0); wb[1][mu] = a2;     wb[1][mv] = R(0);         sb[1] = v;
        wb[2][0] = a3;   wb[2][mu] = a3 * u; wb[2][mv] = a3 * (v + a); sb[2] = v + a;
#pragma unroll
        for (int m = 0; m < 3; m++) {
            R m_ = R(0), p_ = R(0);
#pragma unroll
            for (int mw = 0; mw < 3; mw++) {
                if (mw == 1 && m != mu) continue;   // wb[1][0], wb[1][mv] are zero
                m_ = m_ + rmin(sb[mw], R(0)) * wb[mw][m];
                p_ = p_ + rmax(sb[mw], R(0)) * wb[mw][m];
            }
            bm[m] = m_;
            bp[m] = p_;
        }
    }

    // physical flux f(q) (IXY = 1) or g(q) (IXY = 2), used by the flux form
    template <int IXY>
    __device__ __forceinline__ static void flux(const double* q, const double* /*aux*/,
                                                const PhysParams& pp, double* f) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const double h = q[0], un = q[mu] / h;
        f[0] = q[mu];
        f[mu] = q[mu] * un + 0.5 * pp.grav * h * h;
        f[mv] = q[mv] * un;
    }

    // f(q) and g(q) of one cell with one division (the flux form's update).
    // Never contracted into FMAs: both cells beside an interface evaluate it
    // and must get the same bits.
    __device__ __forceinline__ static void fluxes(const double* q, const double* /*aux*/,
                                                  const PhysParams& pp, double* f, double* g) {
#pragma clang fp contract(off)
        const double h = q[0], rh = flux_rcp(h), u = q[1] * rh, v = q[2] * rh;
        const double p = 0.5 * pp.grav * h * h;
        f[0] = q[1]; f[1] = q[1] * u + p; f[2] = q[2] * u;
        g[0] = q[2]; g[1] = q[1] * v;     g[2] = q[2] * v + p;
    }

    // conservative flux stored in the time-sync registers, operation for
    // operation as applications/clawpack/shallow/2d/radialdam/rpn2cons_update.f:27-46
    template <int IXY>
    __device__ __forceinline__ static void consflux(const double* q, const double* /*aux*/,
                                                    const PhysParams& pp, double* f) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const double hun = q[mu], hut = q[mv], h = q[0];
        const double un = hun / h;
        f[0] = hun;
        f[mu] = (hun * hun) / h + (0.5 * pp.grav) * (h * h);
        f[mv] = un * hut;
    }
};

// ----------------------------------------------------------------- Euler
template <typename R>
struct RPEuler {
    static constexpr int MEQN = 4, MWAVES = 4, MAUX = 0;
    static constexpr bool CONSERVATIVE = true;
    static constexpr bool RPT2_LINEAR = true;   // bm, bp linear in asdq
    // u, v, a, enth, u2v2, g1a2, euv (and in FP32 1 / (2 a), which rpt2 then
    // takes from the normal solve)
    static constexpr int NROE = sizeof(R) == 4 ? 8 : 7;

    // the shear wave carries no density and no normal momentum
    template <int IXY>
    __host__ __device__ static constexpr bool wave_zero(int mw, int m) {
        return mw == 1 && (m == 0 || m == (IXY == 1 ? 1 : 2));
    }

    // What rpn2 computes from one cell alone (n: the sweep's normal
    // direction): sqrt(rho), rho u_n / sqrt(rho), rho u_t / sqrt(rho),
    // (E + p) / sqrt(rho) with p from (rho u)^2 + (rho v)^2, u_n and the sound
    // speed of the entropy fix, whose p sums the momenta in the sweep's order.
    // The step kernel evaluates them once per cell (the same operations, so
    // the same bits).  FP32 takes sqrt(rho), 1 / sqrt(rho) and 1 / rho from one
    // hardware reciprocal square root; FP64 keeps the reference's divisions.
    static constexpr int NCELL = 6;
    template <int IXY>
    __device__ __forceinline__ static void cell(const R* q, const PhysParams& pp, R c[NCELL]) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const R gamma = (R) pp.gamma;
        const R gamma1 = (R) (pp.gamma - 1.0);
        if constexpr (sizeof(R) == 4) {
            const float r = __builtin_amdgcn_rsqf(q[0]), rr = r * r;
            // one pressure (the two summation orders differ only in rounding)
            const float p = gamma1 * (q[3] - 0.5f * (q[1] * q[1] + q[2] * q[2]) * rr);
            c[0] = q[0] * r;
            c[1] = q[mu] * r;
            c[2] = q[mv] * r;
            c[3] = (q[3] + p) * r;
            c[4] = c[1] * r;
            c[5] = __builtin_amdgcn_sqrtf(gamma * p) * r;
            return;
        }
        const R rs = sq(q[0]);
        const R p = gamma1 * (q[3] - dv(R(0.5) * (q[1] * q[1] + q[2] * q[2]), q[0]));
        const R pn = gamma1 * (q[3] - dv(R(0.5) * (q[mu] * q[mu] + q[mv] * q[mv]), q[0]));
        c[0] = rs;
        c[1] = dv(q[mu], rs);
        c[2] = dv(q[mv], rs);
        c[3] = dv(q[3] + p, rs);
        c[4] = dv(q[mu], q[0]);
        c[5] = sq(dv(gamma * pn, q[0]));
    }

    template <int IXY>
    __device__ __forceinline__ static void rpn2(const R* qL, const R* qR, const R* auxL,
                                                const R* auxR, const PhysParams& pp,
                                                R wave[MWAVES][MEQN], R s[MWAVES], R amdq[MEQN],
                                                R apdq[MEQN], R roe[NROE]) {
        R cL[NCELL], cR[NCELL];
        cell<IXY>(qL, pp, cL);
        cell<IXY>(qR, pp, cR);
        rpn2c<IXY>(qL, qR, cL, cR, auxL, auxR, pp, wave, s, amdq, apdq, roe);
    }

    // rpn2 from the two cells' values of `cell`
    template <int IXY>
    __device__ __forceinline__ static void rpn2c(const R* qL, const R* qR, const R* cL,
                                                 const R* cR, const R* /*auxL*/,
                                                 const R* /*auxR*/, const PhysParams& pp,
                                                 R wave[MWAVES][MEQN], R s[MWAVES], R amdq[MEQN],
                                                 R apdq[MEQN], R roe[NROE]) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const R gamma = (R) pp.gamma;
        const R gamma1 = (R) (pp.gamma - 1.0);
        const R rhsqrtl = cL[0];
        const R rhsqrtr = cR[0];
        const R rhsq2 = rhsqrtl + rhsqrtr;
        R u, v, enth, a, g1a2, inv2a = R(0);
        if constexpr (sizeof(R) == 4) {
            // one reciprocal and one reciprocal square root: a = a2s rsq(a2s),
            // gamma1 / a2s = gamma1 rsq(a2s)^2, 1 / (2 a) = rsq(a2s) / 2
            const float rh = __builtin_amdgcn_rcpf(rhsq2);
            u = (cL[1] + cR[1]) * rh;
            v = (cL[2] + cR[2]) * rh;
            enth = (cL[3] + cR[3]) * rh;
        } else {
            u = dv(cL[1] + cR[1], rhsq2);
            v = dv(cL[2] + cR[2], rhsq2);
            enth = dv(cL[3] + cR[3], rhsq2);
        }
        const R u2v2 = u * u + v * v;
        const R a2s = gamma1 * (enth - R(.5) * u2v2);
        if constexpr (sizeof(R) == 4) {
            const float ra = __builtin_amdgcn_rsqf(a2s);
            a = a2s * ra;
            g1a2 = gamma1 * (ra * ra);
            inv2a = 0.5f * ra;
            roe[7] = inv2a;
        } else {
            a = sq(a2s);
            g1a2 = dv(gamma1, a2s);
        }
        const R euv = enth - u2v2;
        roe[0] = u; roe[1] = v; roe[2] = a; roe[3] = enth; roe[4] = u2v2; roe[5] = g1a2; roe[6] = euv;

        const R d0 = qR[0] - qL[0], d1 = qR[mu] - qL[mu], d2 = qR[mv] - qL[mv], d3 = qR[3] - qL[3];
        const R a3 = g1a2 * (euv * d0 + u * d1 + v * d2 - d3);
        const R a2 = d2 - v * d0;
        R a4;
        if constexpr (sizeof(R) == 4) a4 = (d1 + (a - u) * d0 - a * a3) * inv2a;
        else a4 = dv(d1 + (a - u) * d0 - a * a3, R(2) * a);
        const R a1 = d0 - a3 - a4;

        wave[0][0] = a1; wave[0][mu] = a1 * (u - a); wave[0][mv] = a1 * v;
        wave[0][3] = a1 * (enth - u * a); s[0] = u - a;
        wave[1][0] = R(0); wave[1][mu] = R(0); wave[1][mv] = a2;
        wave[1][3] = a2 * v; s[1] = u;
        wave[2][0] = a3; wave[2][mu] = a3 * u; wave[2][mv] = a3 * v;
        wave[2][3] = a3 * R(0.5) * u2v2; s[2] = u;
        wave[3][0] = a4; wave[3][mu] = a4 * (u + a); wave[3][mv] = a4 * v;
        wave[3][3] = a4 * (enth + u * a); s[3] = u + a;

        // entropy fix
        const R s0 = cL[4] - cL[5];   // u_n - c of the left cell
        if (s0 >= R(0) && s[0] > R(0)) {
#pragma unroll
            for (int m = 0; m < 4; m++) amdq[m] = R(0);
        } else {
            const R rho1 = qL[0] + wave[0][0];
            const R rhou1 = qL[mu] + wave[0][mu];
            const R rhov1 = qL[mv] + wave[0][mv];
            const R en1 = qL[3] + wave[0][3];
            const R p1 = gamma1 * (en1 - dv(R(0.5) * (rhou1 * rhou1 + rhov1 * rhov1), rho1));
            const R c1 = sq(dv(gamma * p1, rho1));
            const R s1 = dv(rhou1, rho1) - c1;
            R sfract;
            if (s0 < R(0) && s1 > R(0)) sfract = dv(s0 * (s1 - s[0]), s1 - s0);
            else if (s[0] < R(0)) sfract = s[0];
            else sfract = R(0);
#pragma unroll
            for (int m = 0; m < 4; m++) amdq[m] = sfract * wave[0][m];
            if (!(s[1] >= R(0))) {
#pragma unroll
                for (int m = 0; m < 4; m++) {
                    amdq[m] = amdq[m] + s[1] * wave[1][m];
                    amdq[m] = amdq[m] + s[2] * wave[2][m];
                }
                const R s3 = cR[4] + cR[5];   // u_n + c of the right cell
                const R rho2 = qR[0] - wave[3][0];
                const R rhou2 = qR[mu] - wave[3][mu];
                const R rhov2 = qR[mv] - wave[3][mv];
                const R en2 = qR[3] - wave[3][3];
                const R p2 = gamma1 * (en2 - dv(R(0.5) * (rhou2 * rhou2 + rhov2 * rhov2), rho2));
                const R c2 = sq(dv(gamma * p2, rho2));
                const R s2 = dv(rhou2, rho2) + c2;
                bool add = true;
                if (s2 < R(0) && s3 > R(0)) sfract = dv(s2 * (s3 - s[3]), s3 - s2);
                else if (s[3] < R(0)) sfract = s[3];
                else add = false;
                if (add) {
#pragma unroll
                    for (int m = 0; m < 4; m++) amdq[m] = amdq[m] + sfract * wave[3][m];
                }
            }
        }
#pragma unroll
        for (int m = 0; m < 4; m++) {
            R df = R(0);
#pragma unroll
            for (int mw = 0; mw < 4; mw++)
                if (!wave_zero<IXY>(mw, m)) df = df + s[mw] * wave[mw][m];
            apdq[m] = df - amdq[m];
        }
    }

    template <int IXY>
    __device__ __forceinline__ static void rpt2(const R roe[NROE], const R asdq[MEQN],
                                                const R* /*aux1*/, const R* /*aux2*/,
                                                const R* /*aux3*/, const PhysParams&, R bm[MEQN],
                                                R bp[MEQN]) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const R u = roe[0], v = roe[1], a = roe[2], enth = roe[3], u2v2 = roe[4];
        const R g1a2 = roe[5], euv = roe[6];
        const R a3 = g1a2 * (euv * asdq[0] + u * asdq[mu] + v * asdq[mv] - asdq[3]);
        const R a2 = asdq[mu] - u * asdq[0];
        R a4;
        if constexpr (sizeof(R) == 4) a4 = (asdq[mv] + (a - v) * asdq[0] - a * a3) * roe[7];
        else a4 = dv(asdq[mv] + (a - v) * asdq[0] - a * a3, R(2) * a);
        const R a1 = asdq[0] - a3 - a4;
        R wb[3][4], sb[3];
        wb[0][0] = a1; wb[0][mu] = a1 * u; wb[0][mv] = a1 * (v - a);
        wb[0][3] = a1 * (enth - v * a); sb[0] = v - a;
        wb[1][0] = a3; wb[1][mu] = a3 * u + a2; wb[1][mv] = a3 * v;
        wb[1][3] = a3 * R(0.5) * u2v2 + a2 * u; sb[1] = v;
        wb[2][0] = a4; wb[2][mu] = a4 * u; wb[2][mv] = a4 * (v + a);
        wb[2][3] = a4 * (enth + v * a); sb[2] = v + a;
#pragma unroll
        for (int m = 0; m < 4; m++) {
            R m_ = R(0), p_ = R(0);
#pragma unroll
            for (int mw = 0; mw < 3; mw++) {
                m_ = m_ + rmin(sb[mw], R(0)) * wb[mw][m];
                p_ = p_ + rmax(sb[mw], R(0)) * wb[mw][m];
            }
            bm[m] = m_;
            bp[m] = p_;
        }
    }

    template <int IXY>
    __device__ __forceinline__ static void flux(const double* q, const double* /*aux*/,
                                                const PhysParams& pp, double* f) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const double rho = q[0], un = q[mu] / rho;
        const double p = (pp.gamma - 1.0) * (q[3] - 0.5 * (q[1] * q[1] + q[2] * q[2]) / rho);
        f[0] = q[mu];
        f[mu] = q[mu] * un + p;
        f[mv] = q[mv] * un;
        f[3] = (q[3] + p) * un;
    }

    // f(q) and g(q) of one cell with one division (the flux form's update).
    // Never contracted into FMAs: both cells beside an interface evaluate it
    // and must get the same bits.
    __device__ __forceinline__ static void fluxes(const double* q, const double* /*aux*/,
                                                  const PhysParams& pp, double* f, double* g) {
#pragma clang fp contract(off)
        const double rr = flux_rcp(q[0]), u = q[1] * rr, v = q[2] * rr;
        const double p = (pp.gamma - 1.0) * (q[3] - 0.5 * (q[1] * u + q[2] * v));
        f[0] = q[1]; f[1] = q[1] * u + p; f[2] = q[2] * u; f[3] = (q[3] + p) * u;
        g[0] = q[2]; g[1] = q[1] * v;     g[2] = q[2] * v + p; g[3] = (q[3] + p) * v;
    }

    // conservative flux stored in the time-sync registers (the reference has
    // no Euler rpn2_cons; this is the physical flux, evaluated as the oracle does)
    template <int IXY>
    __device__ __forceinline__ static void consflux(const double* q, const double* /*aux*/,
                                                    const PhysParams& pp, double* f) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const double rho = q[0], un = q[mu] / rho, ut = q[mv] / rho, en = q[3];
        const double p = (pp.gamma - 1.0) * (en - 0.5 * rho * (un * un + ut * ut));
        f[0] = q[mu];
        f[mu] = q[mu] * un + p;
        f[mv] = q[mu] * ut;
        f[3] = un * (en + p);
    }
};

// ----------------------------------------------------------------- advection
// aux(:,:,1) = u, aux(:,:,2) = v.  The normal speed is the right cell's aux
// component IXY; rpt2 takes the transverse velocity from aux2 (bmasdq) and
// aux3 (bpasdq) as clawpack46_rpt2adv.f:79-86.
template <typename R>
struct RPAdvection {
    static constexpr int MEQN = 1, MWAVES = 1, MAUX = 2;
    // colour equation q_t + u q_x + v q_y = 0: the fluctuations are not a flux
    // difference unless the velocity is uniform, so no flux form exists
    static constexpr bool CONSERVATIVE = false;
    static constexpr bool RPT2_LINEAR = true;   // bm, bp linear in asdq
    static constexpr int NROE = 1;

    template <int IXY>
    __device__ __forceinline__ static void rpn2(const R* qL, const R* qR, const R* /*auxL*/,
                                                const R* auxR,
                                                const PhysParams&, R wave[MWAVES][MEQN],
                                                R s[MWAVES], R amdq[MEQN], R apdq[MEQN],
                                                R roe[NROE]) {
        const R delta = qR[0] - qL[0];
        wave[0][0] = delta;
        s[0] = auxR[IXY - 1];
        amdq[0] = R(0) + rmin(s[0], R(0)) * wave[0][0];
        apdq[0] = R(0) + rmax(s[0], R(0)) * wave[0][0];
        roe[0] = R(0);
    }

    template <int IXY>
    __device__ __forceinline__ static void rpt2(const R* /*roe*/, const R asdq[MEQN],
                                                const R* /*aux1*/, const R* aux2, const R* aux3,
                                                const PhysParams&, R bm[MEQN], R bp[MEQN]) {
        constexpr int kv = IXY == 1 ? 1 : 0;   // the transverse velocity component
        bm[0] = rmin(aux2[kv], R(0)) * asdq[0];
        bp[0] = rmax(aux3[kv], R(0)) * asdq[0];
    }

    template <int IXY>
    __device__ __forceinline__ static void flux(const double* q, const double* aux,
                                                const PhysParams&, double* f) {
        f[0] = aux[IXY - 1] * q[0];
    }

    __device__ __forceinline__ static void fluxes(const double* q, const double* aux,
                                                  const PhysParams&, double* f, double* g) {
#pragma clang fp contract(off)
        f[0] = aux[0] * q[0];
        g[0] = aux[1] * q[0];
    }

    // conservative flux stored in the time-sync registers: u_n q with the
    // cell's velocity (applications/clawpack/advection/2d/adv_order3/user_4.6/rpn2cons_update.f:14-26)
    template <int IXY>
    __device__ __forceinline__ static void consflux(const double* q, const double* aux,
                                                    const PhysParams&, double* f) {
        f[0] = aux[IXY - 1] * q[0];
    }
};

}  // namespace hipclaw
