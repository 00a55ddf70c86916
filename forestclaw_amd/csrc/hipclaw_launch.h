// Launch dispatch of the step kernel for one Riemann functor: tile size,
// waves per workgroup, time-sync registers and the compile-time MC limiter
// pick one instantiation of step_tiles.  Shared by the built-in solvers' two
// translation units (FP64 compiled without FMA contraction, FP32-bulk with it)
// and by solvers an application compiles itself (hipclaw_user_solver.h); the
// instantiations take the contraction mode of the unit that includes this.
#pragma once

#include <atomic>
#include <type_traits>

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "hipclaw_step.h"

namespace hipclaw {

// The stream of a step launch and, for a timed launch, two events that
// hipExtLaunchKernelGGL binds to the kernel's own start and end, so their
// elapsed time is the kernel's execution without the gaps around it.
struct StepLaunch {
    hipStream_t st;
    hipEvent_t start, stop;
    StepLaunch(hipStream_t s, hipEvent_t a = nullptr, hipEvent_t b = nullptr) : st(s), start(a), stop(b) {}
};

// Workgroups of `kernel` (block size `block`, static LDS only) that the
// current device holds at once: the grid of a persistent launch.
inline int resident_blocks(const void* kernel, int block) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess) return 0;
    return per_cu * cus;
}

// Launch a persistent step kernel: min(tiles, resident workgroups) of them.
// `resident` caches the occupancy query; each kernel instantiation owns one.
// Host threads may race to fill it: they compute the same value, and the
// atomic makes the store and the loads well defined.
template <typename K>
inline int launch_persistent(K kernel, int block, const StepArgs& a, const StepLaunch& L,
                             std::atomic<int>& resident) {
    const int ntiles = a.nwork * a.ntx * a.nty;
    if (ntiles <= 0) return 0;
    int nres = resident.load(std::memory_order_relaxed);
    if (nres <= 0) {
        nres = resident_blocks((const void*) kernel, block);
        resident.store(nres, std::memory_order_relaxed);
    }
    if (nres <= 0) return (int) hipErrorInvalidDeviceFunction;
    const int nblocks = ntiles < nres ? ntiles : nres;
    if (L.start || L.stop)
        hipExtLaunchKernelGGL(kernel, dim3(nblocks), dim3(block), 0, L.st, L.start, L.stop, 0, a);
    else
        hipLaunchKernelGGL(kernel, dim3(nblocks), dim3(block), 0, L.st, a);
    return (int) hipGetLastError();
}

// Every wave family uses the MC limiter (mthlim 4), so the step may take the
// instantiation with the limiter fixed at compile time.
template <class RP>
inline bool all_mc_limiter(const StepArgs& a) {
    for (int mw = 0; mw < RP::MWAVES; mw++)
        if (a.mthlim[mw] != 4) return false;
    return true;
}

// Whether the functor provides the conservative flux the time-sync registers
// store (consflux<IXY>); without it a solver runs without time sync.
template <class RP, class = void>
struct has_consflux : std::false_type {};
template <class RP>
struct has_consflux<RP, std::void_t<decltype(&RP::template consflux<1>)>> : std::true_type {};

// A step instantiation fits the 160 KiB of LDS a workgroup may hold.
template <class RP, typename R, int T, bool FLUXFORM, bool REGS>
constexpr bool step_fits() {
    return sizeof(TileSmem<T, RP::MEQN, RP::MAUX, R, FLUXFORM, REGS>) + 64 <= 160 * 1024;
}

// One tile shape of one precision.  R = double: the FP64 parity form;
// R = float: the FP32-bulk form (flux form for a conservative solver).
// WITH_REGS = false leaves out the time-sync instantiations.
template <class RP, typename R, int T, int NW, bool WITH_REGS>
inline int launch_step_shape(const StepArgs& a, const StepLaunch& st) {
    // kernels that keep time-sync registers are separate instantiations, so the
    // plain update carries none of their code or register pressure; the MC
    // limiter on every wave family has its own compile-time instantiation
    constexpr bool FF = sizeof(R) == 4 && RP::CONSERVATIVE;
    static std::atomic<int> resident[2][2];   // zero-initialised (static storage)
    if constexpr (!step_fits<RP, R, T, FF, false>()) {
        return (int) hipErrorInvalidValue;
    } else {
        const bool mc = all_mc_limiter<RP>(a);
        if (a.reg) {
            if constexpr (WITH_REGS && step_fits<RP, R, T, FF, true>()) {
                if (mc)
                    return launch_persistent(step_tiles<RP, R, T, NW, FF, true, 4>, NW * 64, a, st, resident[1][1]);
                return launch_persistent(step_tiles<RP, R, T, NW, FF, true>, NW * 64, a, st, resident[0][1]);
            }
            return (int) hipErrorInvalidValue;
        }
        if (mc)
            return launch_persistent(step_tiles<RP, R, T, NW, FF, false, 4>, NW * 64, a, st, resident[1][0]);
        return launch_persistent(step_tiles<RP, R, T, NW, FF, false>, NW * 64, a, st, resident[0][0]);
    }
}

// Waves per 16 x 16 tile by default: four for FP64, two for the FP32-bulk
// form (hipclaw_set_waves and the autotuner choose between 2 and 4).  A sweep
// covers T + 2 = 18 rows, three per wave (19 interfaces each), so two waves
// take three row groups each where four take two, two, one and one and wait
// at the barrier, and the smaller workgroups pack more densely on a CU; FP64
// tiles need the latency hiding of four.  Measured on MI355X (autotuner
// trials, ms per step): shallow water FP32 on the benchmark forest 0.133 with
// two waves against 0.150 with four, the Euler shock-bubble 0.076 against
// 0.094, the 64-run ensemble 0.476 against 0.494.
template <typename R>
constexpr int default_waves_t16() { return sizeof(R) == 8 ? 4 : 2; }

// waves = 0: the default for 16 x 16 tiles; 2, 3 or 4 select the variant.
// T = 32 only where WITH32 (its LDS budget is checked per solver).
template <class RP, typename R, bool WITH_REGS, bool WITH32 = false>
inline int launch_step_tile(int tile, int waves, const StepArgs& a, const StepLaunch& st) {
    if (tile == 16 && waves == 0) waves = default_waves_t16<R>();
    switch (tile) {
    case 8: return launch_step_shape<RP, R, 8, 2, WITH_REGS>(a, st);
    case 16:
        return waves == 2 ? launch_step_shape<RP, R, 16, 2, WITH_REGS>(a, st)
             : waves == 3 ? launch_step_shape<RP, R, 16, 3, WITH_REGS>(a, st)
                          : launch_step_shape<RP, R, 16, 4, WITH_REGS>(a, st);
    case 32:
        if constexpr (WITH32) return launch_step_shape<RP, R, 32, 16, WITH_REGS>(a, st);
        return (int) hipErrorInvalidValue;
    default: return (int) hipErrorInvalidValue;
    }
}

}  // namespace hipclaw
