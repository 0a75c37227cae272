// Internal records shared by the host runtime and the gfx950 kernels.
#pragma once
#include <stdint.h>

namespace kt {

// Step log of a recording trace (the differentiable renderer's forward
// pass): per ray, chunks of K committed steps, each holding the state at
// the chunk's first step (a checkpoint), the link to the ray's previous
// chunk and, per step, the committed step size and the transmittance at
// the step's start.  A ray's first chunk is chunk `pixel`; later chunks
// come from a pool starting at W * H.
struct GradLog {
    double* arena;        // chunks of `stride` doubles
    uint32_t* head;       // per pixel: the ray's last chunk
    uint32_t* nsteps;     // per pixel: committed steps
    double* pk;           // per pixel: pt, kt
    uint32_t* pool_head;  // next free pool chunk
    uint32_t* overflow;   // set when the pool is exhausted
    uint32_t capacity;    // chunks in the arena
    uint32_t K;           // steps per chunk (checkpoint interval)
    uint32_t stride;      // doubles per chunk: 8 + 2 K
    uint32_t reserved;
};
// chunk layout: [0..5] state, [6] previous chunk (bits; ~0 for none),
// [7] unused, [8 + u] h of step u, [8 + K + u] transmittance before step u
static constexpr uint32_t kNoChunk = 0xFFFFFFFFu;

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
    uint32_t long_ray;        // attempts after which a ray raises its wave's priority
    uint32_t isolate;         // 0, or the priority level from which a wave takes no new rays
    // work queue: items are (global tile, pixel within the tile); the
    // tiles of all renders of a launch are numbered consecutively
    uint32_t tile_begin;      // first global tile of this render
    uint32_t n_tiles;         // tiles of this render: tiles_x * tiles_y
    int32_t tile;             // tile edge in pixels (the same for every render of a launch)
    int32_t tiles_x;
    const uint32_t* tile_perm;  // visiting order of the launch's global tiles, or null
    GradLog log;              // recording traces only
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
