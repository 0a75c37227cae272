// Internal records shared by the host runtime and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace kt {

// Everything one render's rays read, derived on the host from kt_config
// (src/renderer/utils.rs:53-114, init.rs:43-73) plus the camera
// (cuda/kernel.cu:88-97).  One record per render; a batched launch holds an
// array of them.  Plain data only: it is copied to the device as is.
struct RenderParams {
    // Kerr geometry (KerrParams, cuda/kerr_math.cuh:3-9)
    double a, M, aa, inv_M;
    double rh_eps;          // r_h + horizon_epsilon
    double disk_inner, disk_outer, T_scale;
    // integrator (CONFIG_INTEGRATOR_*)
    double h0, tol, tol_floor, cutoff, R_esc;
    // pixel -> ray (kernel.cu:112-121)
    double inv_w_2, inv_h_2, aspect, exposure_over_spp;
    double cam[3], fwd[3], rgt[3], up[3], fov_scale;
    // sky (kerr_math.cuh:157-171)
    double sky_spacing, sky_lt, sky_int;
    // disk noise (kerr_trace.cuh:200-214)
    double noise_scale, noise_strength, noise_winding;
    uint64_t noise_seed_mix;  // seed * 0x9E3779B97F4A7C15
    // blackbody LUT fetch (kernel.cu:1-16)
    double wl_step, lut_max_temp;
    // lookup tables in device memory (FP64 and FP32 copies)
    const double* disk_lut;
    const double* bb_lut;     // rgb triplets
    const float* disk_lut_f;
    const float* bb_lut_f;
    // outputs
    float* hdr;               // float4 per pixel
    uint8_t* cls;             // may be null
    uint32_t* attempts_out;   // per-pixel DP5(4) attempts, may be null
    uint32_t* error_flag;     // [0] set once, [1] float bits of the first overflowing T
    int32_t W, H, spp, max_steps, max_attempts, lut_n, noise_detail, disk_root_mode;
    int32_t rhs_exact;        // 1: exact Hamiltonian dr^2/dy
    int32_t reserved;
    // work queue: items are (global tile, pixel within the tile); the
    // tiles of all renders of a launch are numbered consecutively
    uint32_t tile_begin;      // first global tile of this render
    uint32_t n_tiles;         // tiles of this render: tiles_x * tiles_y
    int32_t tile;             // tile edge in pixels (the same for every render of a launch)
    int32_t tiles_x;
    const uint32_t* tile_perm;  // visiting order of the launch's global tiles, or null
};

// Work counters, summed over the launch (one atomic per wave at exit).
struct DevCounters {
    unsigned long long rhs_evals;
    unsigned long long attempts;
    unsigned long long accepted;
    unsigned long long crossings;
    unsigned long long max_attempts_ray;
    unsigned long long queue_head;   // persistent work-queue head
    unsigned long long clock_start;  // earliest wave start (wall clock ticks)
    unsigned long long clock_drain;  // earliest time a wave found the queue empty
    unsigned long long clock_end;    // latest wave exit
};

}  // namespace kt
