// Linear acoustics as an application's own Riemann solver, plugged into the
// step kernel through hipclaw_user_solver.h.  The reference ships it as the
// device solvers of its radial acoustics example
// (applications/cudaclaw/acoustics/2d/radial/user_cuda/rpn2acoustics.cu:30-64,
// rpt2acoustics.cu:6-31), assigned to the vtable's cuda_rpn2 / cuda_rpt2
// (radial_user.cpp:58-62); the arithmetic here follows the CPU solvers
// applications/clawpack/acoustics/2d/rp/clawpack46_rpn2_acoustics.f:39-75 and
// clawpack46_rpt2_acoustics.f:36-61 operation by operation.
//
// q = (p, u, v); constants pp.p = (rho, bulk, cc, zz), the common /cparam/ of
// the acoustics setprob (cc = sqrt(bulk / rho), zz = rho cc).
#include "hipclaw_user_solver.h"

template <typename R>
struct RPAcoustics {
    static constexpr int MEQN = 3, MWAVES = 2, MAUX = 0;
    static constexpr int NROE = 1;
    // f(q) = (bulk u, p / rho, 0) is linear: amdq + apdq = f(qR) - f(qL)
    static constexpr bool CONSERVATIVE = true;

    template <int IXY>
    __device__ __forceinline__ static void rpn2(const R* qL, const R* qR, const R*, const R*,
                                                const hipclaw::PhysParams& pp,
                                                R wave[MWAVES][MEQN], R s[MWAVES], R amdq[MEQN],
                                                R apdq[MEQN], R roe[NROE]) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const R cc = (R) pp.p[2], zz = (R) pp.p[3];
        const R d1 = qR[0] - qL[0];
        const R d2 = qR[mu] - qL[mu];
        const R a1 = hipclaw::dv(-d1 + zz * d2, R(2) * zz);
        const R a2 = hipclaw::dv(d1 + zz * d2, R(2) * zz);
        wave[0][0] = -a1 * zz; wave[0][mu] = a1; wave[0][mv] = R(0); s[0] = -cc;
        wave[1][0] = a2 * zz;  wave[1][mu] = a2; wave[1][mv] = R(0); s[1] = cc;
#pragma unroll
        for (int m = 0; m < MEQN; m++) {
            amdq[m] = s[0] * wave[0][m];
            apdq[m] = s[1] * wave[1][m];
        }
        roe[0] = R(0);
    }

    template <int IXY>
    __device__ __forceinline__ static void rpt2(const R*, const R asdq[MEQN], const R*, const R*,
                                                const R*, const hipclaw::PhysParams& pp,
                                                R bm[MEQN], R bp[MEQN]) {
        constexpr int mu = IXY == 1 ? 1 : 2, mv = IXY == 1 ? 2 : 1;
        const R cc = (R) pp.p[2], zz = (R) pp.p[3];
        const R a1 = hipclaw::dv(-asdq[0] + zz * asdq[mv], R(2) * zz);
        const R a2 = hipclaw::dv(asdq[0] + zz * asdq[mv], R(2) * zz);
        bm[0] = cc * a1 * zz; bm[mu] = R(0); bm[mv] = -cc * a1;
        bp[0] = cc * a2 * zz; bp[mu] = R(0); bp[mv] = cc * a2;
    }

    // the FP32-bulk form's FP64 interface fluxes
    __device__ __forceinline__ static void fluxes(const double* q, const double*,
                                                  const hipclaw::PhysParams& pp, double* f,
                                                  double* g) {
#pragma clang fp contract(off)
        const double rho = pp.p[0], bulk = pp.p[1];
        f[0] = bulk * q[1]; f[1] = q[0] / rho; f[2] = 0.0;
        g[0] = bulk * q[2]; g[1] = 0.0;        g[2] = q[0] / rho;
    }
};

HIPCLAW_USER_SOLVER(acoustics, RPAcoustics)
