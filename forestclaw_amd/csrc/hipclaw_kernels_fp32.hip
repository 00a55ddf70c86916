// FP32-bulk patch step: Riemann solves, limiter and transverse splitting in
// FP32, conservative FP64 interface fluxes and FP64 q storage.  Compiled with
// FMA contraction allowed (this form is measured against the FP64 oracle with
// a stated error tolerance, not bit for bit).
#include "hipclaw_internal.h"

namespace hipclaw {

int launch_step_fp32(int solver, int tile, int waves, const StepArgs& a, const StepLaunch& st) {
    switch (solver) {
    case 0: return launch_step_tile<RPAdvection<float>, float, true, true>(tile, waves, a, st);
    case 1: return launch_step_tile<RPShallow<float>, float, true, true>(tile, waves, a, st);
    case 2:  // meqn = 4 at T = 32: not instantiated (step_supported)
        return launch_step_tile<RPEuler<float>, float, true>(tile, waves, a, st);
    default: return (int) hipErrorInvalidValue;
    }
}

bool step_supported(int solver, int precision, int tile) {
    if (tile != 8 && tile != 16 && tile != 32) return false;
    // T = 32 fits the 160 KB of LDS for meqn = 3 (FP64) and meqn <= 3 (FP32)
    if (tile == 32) return precision == 0 ? solver == 1 : solver != 2;
    return solver >= 0 && solver <= 2;
}

int default_waves_fp32(int /*solver*/, int tile) { return tile == 16 ? default_waves_t16<float>() : 0; }

}  // namespace hipclaw
