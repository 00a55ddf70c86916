// FP32-bulk patch step: Riemann solves, limiter and transverse splitting in
// FP32, conservative FP64 interface fluxes and FP64 q storage.  Compiled with
// FMA contraction allowed (this form is measured against the FP64 oracle with
// a stated error tolerance, not bit for bit).
#include "hipclaw_internal.h"

namespace hipclaw {

template <class RP, int T, int NW>
static int launch_one(const StepArgs& a, const StepLaunch& st) {
    // kernels that keep time-sync registers are separate instantiations, so the
    // plain update carries none of their code or register pressure; the MC
    // limiter on every wave family has its own compile-time instantiation
    static std::atomic<int> resident[2][2];   // zero-initialised (static storage)
    constexpr bool C = RP::CONSERVATIVE;
    if (all_mc_limiter<RP>(a)) {
        if (a.reg)
            return launch_persistent(step_tiles<RP, float, T, NW, C, true, 4>, NW * 64, a, st, resident[1][1]);
        return launch_persistent(step_tiles<RP, float, T, NW, C, false, 4>, NW * 64, a, st, resident[1][0]);
    }
    if (a.reg)
        return launch_persistent(step_tiles<RP, float, T, NW, C, true>, NW * 64, a, st, resident[0][1]);
    return launch_persistent(step_tiles<RP, float, T, NW, C, false>, NW * 64, a, st, resident[0][0]);
}

// Waves per 16 x 16 tile by default (hipclaw_set_waves and the autotuner
// choose between 2 and 4).  A sweep covers T + 2 = 18 rows, three per wave
// (19 interfaces each), so two waves take three row groups each where four
// take two, two, one and one and wait at the barrier, and the smaller
// workgroups pack more densely on a CU.  Measured on MI355X (autotuner
// trials, ms per step): shallow water on the benchmark forest 0.133 with
// two waves against 0.150 with four, the Euler shock-bubble 0.076 against
// 0.094, the 64-run ensemble 0.476 against 0.494.
template <class RP>
constexpr int waves_t16() { return 2; }

// waves = 0: waves_t16 for 16 x 16 tiles; 2 or 4 select the variant.
template <class RP, bool WITH32 = true>
static int launch_tile(int tile, int waves, const StepArgs& a, const StepLaunch& st) {
    if (tile == 16 && waves == 0) waves = waves_t16<RP>();
    switch (tile) {
    case 8: return launch_one<RP, 8, 2>(a, st);
    case 16: return waves == 2 ? launch_one<RP, 16, 2>(a, st) : launch_one<RP, 16, 4>(a, st);
    case 32:
        if constexpr (WITH32) return launch_one<RP, 32, 16>(a, st);
        return (int) hipErrorInvalidValue;
    default: return (int) hipErrorInvalidValue;
    }
}

int launch_step_fp32(int solver, int tile, int waves, const StepArgs& a, const StepLaunch& st) {
    switch (solver) {
    case 0: return launch_tile<RPAdvection<float>>(tile, waves, a, st);
    case 1: return launch_tile<RPShallow<float>>(tile, waves, a, st);
    case 2:  // meqn = 4 at T = 32 needs more LDS than a workgroup may hold
        return launch_tile<RPEuler<float>, false>(tile, waves, a, st);
    default: return (int) hipErrorInvalidValue;
    }
}

bool step_supported(int solver, int precision, int tile) {
    if (tile != 8 && tile != 16 && tile != 32) return false;
    // T = 32 fits the 160 KB of LDS for meqn = 3 (FP64) and meqn <= 3 (FP32)
    if (tile == 32) return precision == 0 ? solver == 1 : solver != 2;
    return solver >= 0 && solver <= 2;
}

int default_waves_fp32(int solver, int tile) {
    if (tile != 16) return 0;
    switch (solver) {
    case 0: return waves_t16<RPAdvection<float>>();
    case 1: return waves_t16<RPShallow<float>>();
    default: return waves_t16<RPEuler<float>>();
    }
}

}  // namespace hipclaw
