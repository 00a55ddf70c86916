// Header-only path for an application's own Riemann solver.
//
// Write the solver as a functor template over the arithmetic type R (double
// for the FP64 form, float for the FP32-bulk form):
//
//   template <typename R> struct MySolver {
//       static constexpr int MEQN = ..., MWAVES = ... (<= 10), MAUX = ...;
//       static constexpr int NROE = ...;         // values rpn2 hands to rpt2
//       static constexpr bool CONSERVATIVE = ...;// amdq + apdq = f(qR) - f(qL)
//
//       // one interface between cell qL (left / below) and cell qR; IXY = 1
//       // for an x-interface, 2 for a y-interface
//       template <int IXY> __device__ static void rpn2(
//           const R* qL, const R* qR, const R* auxL, const R* auxR,
//           const hipclaw::PhysParams& pp, R wave[MWAVES][MEQN], R s[MWAVES],
//           R amdq[MEQN], R apdq[MEQN], R roe[NROE]);
//       // split asdq (amdq or apdq of that interface) in the transverse
//       // direction; aux1, aux2, aux3: the aux values of the three cells of
//       // the transverse line through the cell asdq goes into (rows j-1, j,
//       // j+1 for IXY = 1, columns i-1, i, i+1 for IXY = 2)
//       template <int IXY> __device__ static void rpt2(
//           const R roe[NROE], const R asdq[MEQN], const R* aux1, const R* aux2,
//           const R* aux3, const hipclaw::PhysParams& pp, R bmasdq[MEQN],
//           R bpasdq[MEQN]);
//
//       // only when CONSERVATIVE: f(q) and g(q) of one cell in FP64 (the
//       // FP32-bulk form's shared interface fluxes; write them under
//       // `#pragma clang fp contract(off)` so both cells of an interface get
//       // the same bits)
//       __device__ static void fluxes(const double* q, const double* aux,
//                                     const hipclaw::PhysParams& pp, double* f, double* g);
//       // optional: the conservative flux f(q) (IXY = 1) or g(q) kept in the
//       // time-sync registers; without it the solver runs without time sync
//       template <int IXY> __device__ static void consflux(
//           const double* q, const double* aux, const hipclaw::PhysParams& pp, double* f);
//   };
//
// pp.p[0..7] holds the solver's own constants (hipclaw_config_t.rp_params).
// The arguments and their meaning are those of the reference's device solver
// types cudaclaw_cuda_rpn2_t / cudaclaw_cuda_rpt2_t
// (src/solvers/fc2d_cudaclaw/fc2d_cudaclaw_cuda.h:57-67) with the waves as
// wave[mw][m] and the transverse split of one fluctuation per call.
//
// Then, in one .hip file:
//
//   #include "hipclaw_user_solver.h"
//   ... the functor ...
//   HIPCLAW_USER_SOLVER(my_solver, MySolver)
//
// and compile it twice for gfx950: once with -DHIPCLAW_USER_PRECISION=64
// -ffp-contract=off (the FP64 form, bit-faithful to a CPU solver written in
// the same operation order) and once with -DHIPCLAW_USER_PRECISION=32 (the
// FP32-bulk form: the flux form for a conservative solver, FP32 fluctuations
// on FP64 q otherwise; leave this object out to offer FP64 only); both with
// -Xclang -target-feature -Xclang -packed-fp32-ops.  Link the objects into one
// shared library.  It exports
//
//   const hipclaw_solver_t *my_solver_solver(void);
//
// whose result goes to hipclaw_register_solver (or fc2d_hipclaw_vtable_t's
// rp_solver).
#pragma once

#include "../../include/hipclaw_solver.h"
#include "hipclaw_launch.h"

#ifndef HIPCLAW_USER_PRECISION
#error "compile with -DHIPCLAW_USER_PRECISION=64 (FP64 form) or 32 (FP32-bulk form)"
#endif

namespace hipclaw {

template <class RP, typename R>
inline int user_launch(int tile, int waves, const void* args, void* st, void* e0, void* e1) {
    static_assert(RP::MWAVES >= 1 && RP::MWAVES <= 10, "mwaves must be 1..10");
    constexpr bool regs = sizeof(R) == 4 && RP::CONSERVATIVE ? true : has_consflux<RP>::value;
    return launch_step_tile<RP, R, regs>(tile, waves, *static_cast<const StepArgs*>(args),
                                         StepLaunch((hipStream_t) st, (hipEvent_t) e0, (hipEvent_t) e1));
}

template <class RP>
constexpr int user_tiles() {
    return (step_fits<RP, double, 8, false, false>() ? 8 : 0) |
           (step_fits<RP, double, 16, false, false>() ? 16 : 0);
}

}  // namespace hipclaw

#if HIPCLAW_USER_PRECISION == 64
#define HIPCLAW_USER_SOLVER(NAME, FUNCTOR)                                                          \
    extern "C" int NAME##_launch_fp64(int tile, int waves, const void* args, void* st, void* e0,   \
                                      void* e1) {                                                  \
        return hipclaw::user_launch<FUNCTOR<double>, double>(tile, waves, args, st, e0, e1);       \
    }                                                                                              \
    extern "C" __attribute__((weak)) int NAME##_launch_fp32(int, int, const void*, void*, void*,   \
                                                            void*);                                \
    extern "C" const hipclaw_solver_t* NAME##_solver(void) {                                       \
        using RPD = FUNCTOR<double>;                                                               \
        static const hipclaw_solver_t d = {                                                        \
            HIPCLAW_SOLVER_ABI, (int) sizeof(hipclaw::StepArgs), #NAME, RPD::MEQN, RPD::MWAVES,    \
            RPD::MAUX, RPD::CONSERVATIVE ? 1 : 0, hipclaw::has_consflux<RPD>::value ? 1 : 0,       \
            hipclaw::user_tiles<RPD>(), NAME##_launch_fp64, NAME##_launch_fp32};                   \
        return &d;                                                                                 \
    }
#elif HIPCLAW_USER_PRECISION == 32
#define HIPCLAW_USER_SOLVER(NAME, FUNCTOR)                                                          \
    extern "C" int NAME##_launch_fp32(int tile, int waves, const void* args, void* st, void* e0,   \
                                      void* e1) {                                                  \
        return hipclaw::user_launch<FUNCTOR<float>, float>(tile, waves, args, st, e0, e1);         \
    }
#else
#error "HIPCLAW_USER_PRECISION must be 64 or 32"
#endif
