// Mixed-precision tracer (FP32 stage bulk, FP64 r-solve / error norm /
// controller / disk root / accumulators).  Not built yet: the launcher
// reports hipErrorNotSupported so a mixed-precision request fails loudly.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"

namespace kt {
hipError_t launch_trace_mixed(const RenderParams*, int, uint64_t, int, DevCounters*, int, hipStream_t) {
    return hipErrorNotSupported;
}
}  // namespace kt
