// FP16 storage of (f - w) scaled by a power of two; FP32 arithmetic; FP32 cotangents.
#include "launch.h"

namespace lbm {
const Runner* runner_f16s() {
    static const RunnerImpl<P_F16S, P_F32, true> r;
    return &r;
}
const ThermalRunner* thermal_runner_f16s() {
    static const ThermalRunnerImpl<P_F16S, P_F32> r;
    return &r;
}
}  // namespace lbm
