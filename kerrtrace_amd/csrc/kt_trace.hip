// Persistent FP64 geodesic tracer for gfx950 (the loop is in
// kt_trace_body.hpp; every operation of a ray is binary64).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"
#include "kt_math.hpp"
#include "kt_trace_body.hpp"
#include "kt_tracer_common.hpp"

namespace kt {

template <bool BATCH, bool RECORD, int RHS>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void trace_fp64_persistent(
    const RenderParams* __restrict__ Ps, int n_renders, uint64_t total_items, int refill_threshold,
    DevCounters* __restrict__ counters) {
    trace_loop<BATCH, RECORD, false, RHS>(Ps, n_renders, total_items, refill_threshold, counters);
}

__global__ void reset_counters(DevCounters* c, uint32_t* error_flag) {
    if (threadIdx.x == 0) {
        c->rhs_evals = 0;
        c->attempts = 0;
        c->accepted = 0;
        c->crossings = 0;
        c->max_attempts_ray = 0;
        c->queue_head = 0;
        c->clock_start = ~0ull;
        c->clock_drain = ~0ull;
        c->clock_end = 0;
        if (error_flag) {
            error_flag[0] = 0;
            error_flag[1] = 0;
        }
    }
}

// Host-side launchers (called by the runtime in kt_host.cpp).
hipError_t launch_reset(DevCounters* c, uint32_t* error_flag, hipStream_t stream) {
    hipLaunchKernelGGL(reset_counters, dim3(1), dim3(64), 0, stream, c, error_flag);
    return hipGetLastError();
}

// d_params holds n_renders records sorted by ray_begin (one for a single
// render).  rhs_exact: the single render's right-hand-side form (a batch
// reads each render's own).
hipError_t launch_trace_fp64(const RenderParams* d_params, int n_renders, uint64_t total_items, int refill_threshold,
                             DevCounters* d_counters, int blocks, bool rhs_exact, hipStream_t stream) {
    if (n_renders > 1)
        hipLaunchKernelGGL((trace_fp64_persistent<true, false, kRhsPerRender>), dim3(blocks), dim3(256), 0, stream,
                           d_params, n_renders, total_items, refill_threshold, d_counters);
    else if (rhs_exact)
        hipLaunchKernelGGL((trace_fp64_persistent<false, false, kRhsExact>), dim3(blocks), dim3(256), 0, stream,
                           d_params, 1, total_items, refill_threshold, d_counters);
    else
        hipLaunchKernelGGL((trace_fp64_persistent<false, false, kRhsAsWritten>), dim3(blocks), dim3(256), 0, stream,
                           d_params, 1, total_items, refill_threshold, d_counters);
    return hipGetLastError();
}

// The differentiable renderer's forward pass (one render, step log on; the
// differentiable renderer takes only the exact right-hand side).
hipError_t launch_trace_fp64_record(const RenderParams* d_params, uint64_t total_items, int refill_threshold,
                                    DevCounters* d_counters, int blocks, hipStream_t stream) {
    hipLaunchKernelGGL((trace_fp64_persistent<false, true, kRhsExact>), dim3(blocks), dim3(256), 0, stream, d_params,
                       1, total_items, refill_threshold, d_counters);
    return hipGetLastError();
}

}  // namespace kt
