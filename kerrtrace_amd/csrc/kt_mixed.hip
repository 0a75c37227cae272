// Persistent mixed-precision geodesic tracer for gfx950: the loop of
// kt_trace_body.hpp with MIXED on.  The right-hand side's bulk runs in
// FP32 (rhs32: FP64 radius solve, FP32 factored derivatives, FP64 result
// assembled as p - corr, kp); the stage inputs, the 5th-order solution, the
// embedded error norm, the step controller, the dense-output disk root, the
// state and the shading stay FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"
#include "kt_math.hpp"
#include "kt_trace_body.hpp"
#include "kt_tracer_common.hpp"

namespace kt {

// RECORD: the differentiable renderer's mixed-precision forward pass (step
// log on).
template <bool BATCH, bool RECORD, int RHS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void trace_mixed_persistent(
    const RenderParams* __restrict__ Ps, int n_renders, uint64_t total_items, int refill_threshold,
    DevCounters* __restrict__ counters) {
    trace_loop<BATCH, RECORD, true, RHS>(Ps, n_renders, total_items, refill_threshold, counters);
}

// rhs_exact: the single render's right-hand-side form (a batch reads each
// render's own).
hipError_t launch_trace_mixed(const RenderParams* d_params, int n_renders, uint64_t total_items, int refill_threshold,
                              DevCounters* d_counters, int blocks, bool rhs_exact, hipStream_t stream) {
    if (n_renders > 1)
        hipLaunchKernelGGL((trace_mixed_persistent<true, false, kRhsPerRender>), dim3(blocks), dim3(256), 0, stream,
                           d_params, n_renders, total_items, refill_threshold, d_counters);
    else if (rhs_exact)
        hipLaunchKernelGGL((trace_mixed_persistent<false, false, kRhsExact>), dim3(blocks), dim3(256), 0, stream,
                           d_params, 1, total_items, refill_threshold, d_counters);
    else
        hipLaunchKernelGGL((trace_mixed_persistent<false, false, kRhsAsWritten>), dim3(blocks), dim3(256), 0, stream,
                           d_params, 1, total_items, refill_threshold, d_counters);
    return hipGetLastError();
}

// exact right-hand side only, as launch_trace_fp64_record
hipError_t launch_trace_mixed_record(const RenderParams* d_params, uint64_t total_items, int refill_threshold,
                                     DevCounters* d_counters, int blocks, hipStream_t stream) {
    hipLaunchKernelGGL((trace_mixed_persistent<false, true, kRhsExact>), dim3(blocks), dim3(256), 0, stream, d_params,
                       1, total_items, refill_threshold, d_counters);
    return hipGetLastError();
}

}  // namespace kt
