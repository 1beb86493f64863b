// Double-precision storage and arithmetic.  Built with -ffp-contract=off: together
// with IEEE division this reproduces the reference's FP64 results bit for bit.
#include "launch.h"

namespace lbm {
const Runner* runner_f64() {
    static const RunnerImpl<P_F64, P_F64, true> r;
    return &r;
}
const ThermalRunner* thermal_runner_f64() {
    static const ThermalRunnerImpl<P_F64, P_F64> r;
    return &r;
}
}  // namespace lbm
