// FP16 storage of (f - w) scaled by a power of two; FP32 arithmetic; cotangents stored
// in FP16 with one dynamic power-of-two exponent for the whole field.
#include "launch.h"

namespace lbm {
const Runner* runner_f16s_l16() {
    static const RunnerImpl<P_F16S, P_F16S, true> r;
    return &r;
}
}  // namespace lbm
