// Single-precision storage and arithmetic.
#include "launch.h"

namespace lbm {
const Runner* runner_f32() {
    static const RunnerImpl<P_F32, P_F32, true> r;
    return &r;
}
const ThermalRunner* thermal_runner_f32() {
    static const ThermalRunnerImpl<P_F32, P_F32> r;
    return &r;
}
}  // namespace lbm
