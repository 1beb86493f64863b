// BF16 storage of (f - w); FP32 arithmetic; FP32 cotangents.
#include "launch.h"

namespace lbm {
const Runner* runner_bf16s() {
    static const RunnerImpl<P_BF16S, P_F32, true> r;
    return &r;
}
}  // namespace lbm
