// Ten independent colour equations as an application's own Riemann solver:
// tracer k is carried by the velocity c_k (u, v), c_k = (10 - k) / 10, with
// aux = (u, v) as the advection solver's (clawpack46_rpn2adv.f:28-58,
// clawpack46_rpt2adv.f:79-86).  Wave family k holds tracer k only, so the
// ten families take ten speeds and, per family, their own limiter: the
// solver exercises the step kernel at its largest wave count
// (FC2D_CUDACLAW_MWAVES = 10, src/solvers/fc2d_cudaclaw/fc2d_cudaclaw_cuda.h:42).
// The fluctuations sum over every wave, as rpn2adv's do.
#include "hipclaw_user_solver.h"

template <typename R>
struct RPTracers {
    static constexpr int MEQN = 10, MWAVES = 10, MAUX = 2;
    static constexpr int NROE = 1;
    static constexpr bool CONSERVATIVE = false;   // colour equations, as advection

    __device__ __forceinline__ static R scale(int k) {
        return (R) ((double) (MEQN - k) / (double) MEQN);
    }

    template <int IXY>
    __device__ __forceinline__ static void rpn2(const R* qL, const R* qR, const R*, const R* auxR,
                                                const hipclaw::PhysParams&,
                                                R wave[MWAVES][MEQN], R s[MWAVES], R amdq[MEQN],
                                                R apdq[MEQN], R roe[NROE]) {
#pragma unroll
        for (int mw = 0; mw < MWAVES; mw++) {
#pragma unroll
            for (int m = 0; m < MEQN; m++) wave[mw][m] = m == mw ? qR[m] - qL[m] : R(0);
            s[mw] = scale(mw) * auxR[IXY - 1];
        }
#pragma unroll
        for (int m = 0; m < MEQN; m++) {
            R am = R(0), ap = R(0);
#pragma unroll
            for (int mw = 0; mw < MWAVES; mw++) {
                am = am + hipclaw::rmin(s[mw], R(0)) * wave[mw][m];
                ap = ap + hipclaw::rmax(s[mw], R(0)) * wave[mw][m];
            }
            amdq[m] = am;
            apdq[m] = ap;
        }
        roe[0] = R(0);
    }

    template <int IXY>
    __device__ __forceinline__ static void rpt2(const R*, const R asdq[MEQN], const R*,
                                                const R* aux2, const R* aux3,
                                                const hipclaw::PhysParams&, R bm[MEQN],
                                                R bp[MEQN]) {
        constexpr int kv = IXY == 1 ? 1 : 0;   // the transverse velocity component
#pragma unroll
        for (int m = 0; m < MEQN; m++) {
            bm[m] = hipclaw::rmin(scale(m) * aux2[kv], R(0)) * asdq[m];
            bp[m] = hipclaw::rmax(scale(m) * aux3[kv], R(0)) * asdq[m];
        }
    }
};

HIPCLAW_USER_SOLVER(tracers, RPTracers)
