/*
 * kerrtrace.h -- C ABI of the MI355X (gfx950) Kerr geodesic renderer.
 *
 * The reference (DamoyY/KerrTrace) has no C ABI of its own: its Rust host
 * (`CudaRenderer`, src/renderer/mod.rs:52-293) drives cudarc directly.  The
 * entry points below replace, one for one, what that host does, so a Rust
 * (or any FFI) host binds this library instead of cudarc:
 *
 *   kt_config            <- Config tree, src/main.rs:14-122 (+ config.yaml:1-67)
 *   kt_config_defaults   <- config.yaml:1-67 default values
 *   kt_create            <- CudaRenderer::new, src/renderer/mod.rs:76-117
 *   kt_submit_render     <- CudaRenderer::submit_render, mod.rs:119-218
 *   kt_ready_frame       <- CudaRenderer::ready_frame, mod.rs:220-250
 *   kt_finish_frame      <- CudaRenderer::finish_frame, mod.rs:252-260
 *   kt_destroy           <- Drop of CudaRenderer (mod.rs:52-74 fields)
 *   kt_trace             <- launch of trace_kernel, cuda/kernel.cu:88-142
 *                           (mod.rs:153-173)
 *   kt_post_process      <- launches of bloom_horizontal + post_process,
 *                           cuda/kernel.cu:144-251 (mod.rs:177-206)
 *
 * Additions with no reference counterpart (precision modes, batched
 * launch, adjoint) are marked as such below.
 *
 * Conventions: every function returns a kt_status (0 = ok, < 0 = error);
 * no C++ exception or abort crosses the ABI.  The message of the last error
 * is read with kt_last_error.  Device pointers are plain `void*`/typed
 * pointers into HIP device memory; `stream` is a hipStream_t passed as
 * `void*` (NULL = the null stream).  A renderer handle is owned by one
 * thread and is not re-entrant (the reference's winit loop is single
 * threaded, src/main.rs:205-206).
 */
#ifndef KERRTRACE_H
#define KERRTRACE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t kt_status;
enum {
    KT_OK = 0,
    KT_ERR_INVALID_CONFIG = -1,   /* build_kerr_params / validate_bloom_settings errors */
    KT_ERR_DEVICE = -2,           /* HIP runtime error (allocation, copy, sync) */
    KT_ERR_LAUNCH = -3,           /* kernel launch failure */
    KT_ERR_LUT_OVERFLOW = -4,     /* colour temperature above lut_max_temp (kernel.cu:3-10, mod.rs:236-241) */
    KT_ERR_QUEUE_ORDER = -5,      /* finish_frame out of FIFO order (mod.rs:252-256) */
    KT_ERR_INVALID_ARGUMENT = -6, /* null pointer, size mismatch, bad index */
    KT_ERR_CAPACITY = -7          /* request larger than a fixed internal limit */
};

/* Precision of the tracer.  The reference traces in FP32 with fast math
 * (config.yaml:40, src/renderer/init.rs:114); both modes here are new. */
enum {
    KT_PRECISION_FP64 = 0,  /* every operation of the ray in FP64 */
    KT_PRECISION_MIXED = 1  /* right-hand-side bulk in FP32; r-solve, stage
                               algebra, error norm, step controller, disk
                               root, state and affine accumulators in FP64
                               (renders below mixed_min_tolerance: FP64) */
};

/* Equatorial-crossing root (cuda/kerr_trace.cuh:138-174). */
enum {
    KT_DISK_ROOT_SECANT_REF = 0, /* as written: the Newton loop at lines 147-168
                                    updates a shadowing local, so the crossing
                                    parameter is the clamped secant y0/(y0-y1) */
    KT_DISK_ROOT_TRUE = 1        /* safeguarded Newton/bisection on the DOPRI5
                                    dense-output interpolant, in FP64 */
};

/* Right-hand side of the ray equations (cuda/kerr_math.cuh:82-156). */
enum {
    KT_RHS_REFERENCE = 0,  /* as written: dr^2/dy = y (1 + (u + 4a^2) / (2s)),
                              which halves the u/s term of the exact
                              derivative (kerr_math.cuh:97) */
    KT_RHS_HAMILTONIAN = 1 /* exact gradient of the Kerr-Schild Hamiltonian:
                              dr^2/dy = y (1 + (u + 2a^2) / s) */
};

/* Per-pixel hit-class byte written by kt_trace (new; the reference only
 * writes colour).  Bits 0-1: terminal reason; higher bits: flags. */
enum {
    KT_END_OTHER = 0,       /* step budget or transmittance cutoff exhausted,
                               or the state became non-finite */
    KT_END_HORIZON = 1,     /* r < r_h + horizon_epsilon (kerr_trace.cuh:135-137) */
    KT_END_ESCAPE = 2,      /* r > escape_radius (kerr_trace.cuh:229-237) */
    KT_FLAG_DISK = 4,       /* at least one disk emission was composited */
    KT_FLAG_SKY_GRID = 8,   /* the escape direction fell on a sky grid line */
    KT_FLAG_LUT_OVERFLOW = 16, /* a disk temperature exceeded lut_max_temp */
    KT_DISK_COUNT_SHIFT = 5   /* bits 5-7: disk contributions, saturating at 7 */
};

/*
 * Renderer configuration.  Field for field a superset of the reference's
 * Config tree (src/main.rs:14-122) restricted to what reaches the render
 * path; floating-point knobs are double here (the reference stores f32).
 * Knobs marked "ext" have no reference counterpart.
 */
typedef struct kt_config {
    /* window: main.rs:27-31 */
    uint32_t width;
    uint32_t height;
    /* kernel: main.rs:90-97 */
    uint32_t spp;
    uint32_t sky_grid_divisions;       /* kernel.sky.grid_divisions */
    double exposure_scale;
    double sky_line_thickness;
    double sky_intensity;
    double spin;                       /* kernel.black_hole.spin (a*M, dimensional) */
    double mass;                       /* kernel.black_hole.mass */
    double disk_outer_radius;          /* kernel.disk.outer_radius */
    double disk_temperature_scale;     /* kernel.disk.temperature_scale */
    double initial_step;               /* kernel.integrator.* : main.rs:114-122 */
    double tolerance;
    uint32_t max_steps;
    uint32_t max_attempts;
    double transmittance_cutoff;
    double horizon_epsilon;
    double escape_radius;
    /* blackbody: main.rs:67-73 */
    uint32_t lut_size;
    uint32_t reserved0;
    double lut_max_temp;
    double wavelength_start;
    double wavelength_end;
    double wavelength_step;
    /* disk_noise: main.rs:79-85 */
    int32_t noise_enabled;
    uint32_t noise_detail;
    double noise_scale;
    double noise_strength;
    double noise_winding;
    uint64_t noise_seed;               /* ext: lattice-hash seed; 0 == reference hash */
    /* bloom: main.rs:56-60 */
    int32_t bloom_enabled;
    int32_t reserved1;
    double bloom_intensity;
    double bloom_radius;               /* sigma = radius / 3 (mod.rs:141) */
    double bloom_threshold;            /* ext: bright-pass max(c - thr, 0); 0 == reference */
    /* extensions */
    double tol_floor;                  /* ext: tol(H) = tol * max(5 - 5 t, tol_floor); 0 == reference */
    double disk_peak_temperature;      /* ext: > 0 sets temperature_scale = peak / max(disk LUT) */
    double mdot_scale;                 /* ext: temperature_scale *= mdot_scale^(1/4); 1 == reference */
    int32_t disk_root_mode;            /* ext: KT_DISK_ROOT_* */
    int32_t precision;                 /* ext: KT_PRECISION_* */
    int32_t rhs_mode;                  /* ext: KT_RHS_* */
    int32_t reserved2;
    double mixed_min_tolerance;        /* ext, KT_PRECISION_MIXED: a render whose
                                          tolerance is below this traces in FP64
                                          (FP32 rounding of the right-hand side
                                          would reach the error tolerance);
                                          default 5e-8 */
} kt_config;

/* Pinhole camera in FP64.  The reference passes 12 f32 scalars plus
 * fov_scale = tan(fov/2) (cuda/kernel.cu:88-97, src/app.rs:163). */
typedef struct kt_camera {
    double pos[3];
    double fwd[3];
    double rgt[3];
    double up[3];
    double fov_scale;
} kt_camera;

/* Host-derived quantities (src/renderer/utils.rs:53-114, 172-201). */
typedef struct kt_derived {
    double rh;            /* M + sqrt(M^2 - a^2) */
    double disk_inner;    /* prograde ISCO of |a|/M */
    double disk_outer;
    double temperature_scale; /* after peak normalisation and mdot scaling */
    double disk_lut_max;  /* max_i f^(1/4)[i] */
    double bloom_sigma;
    int32_t bloom_radius_int; /* ceil(bloom.radius) (init.rs:58-66) */
    int32_t bloom_active;
} kt_derived;

/* Work counters of the last kt_trace (new; used for the VALU roofline). */
typedef struct kt_trace_stats {
    uint64_t rays;
    uint64_t rhs_evals;     /* evaluations of the Hamiltonian right-hand side */
    uint64_t attempts;      /* DP5(4) step attempts */
    uint64_t accepted;      /* accepted (not forced) steps */
    uint64_t crossings;     /* equatorial crossings handled */
    uint64_t max_attempts_ray; /* largest attempt count of any single ray */
    uint32_t lut_overflow;  /* 1 if a disk temperature exceeded lut_max_temp */
    float overflow_temperature; /* the first such temperature (kernel.cu:5-8) */
    double kernel_us;       /* first wave start to last wave exit (device clock) */
    double tail_us;         /* work queue empty to last wave exit */
    double kernel_event_ms; /* HIP events recorded around the tracer launch alone */
} kt_trace_stats;

/* Tuning knobs of the persistent tracer (new).  Zero means "default". */
typedef struct kt_launch_opts {
    int32_t waves_per_cu;   /* resident waves per CU of the persistent grid */
    int32_t refill_threshold; /* refill a wave once this many lanes are idle */
    int32_t queue_chunk;    /* pixel tile edge used to order the work queue */
    int32_t cost_order;     /* 0 or 1: visit tiles hardest-first (predicted
                               from the rays' closeness to the photon-orbit
                               critical curve); 2: row-major order */
    int32_t long_ray_attempts; /* a wave holding a ray past this many attempts
                               raises its issue priority (3x, 8x: higher) */
    int32_t isolate_level;  /* 0: off; 1-3: a wave whose priority level (above)
                               reaches this takes no new rays until its long
                               ray ends, so the ray's serial chain is not
                               slowed by wave-mates starting and shading */
} kt_launch_opts;

typedef struct kt_renderer kt_renderer;

typedef struct kt_frame_view {
    int32_t index;            /* FrameView.index (mod.rs:20-23) */
    uint32_t width;
    uint32_t height;
    uint32_t reserved;
    const uint32_t* data;     /* 0x00RRGGBB, borrowed until kt_finish_frame */
} kt_frame_view;

/* ---- configuration ---------------------------------------------------- */
kt_status kt_config_defaults(kt_config* cfg);
kt_status kt_validate_config(const kt_config* cfg, kt_derived* out);
kt_status kt_camera_from_yaw_pitch(const double pos[3], double yaw_deg,
                                   double pitch_deg, double fov_deg,
                                   kt_camera* out); /* src/math.rs:3-14, app.rs:163 */
kt_status kt_camera_observer(double r_obs, double inclination_deg,
                             double fov_deg, kt_camera* out);
kt_status kt_last_error(const kt_renderer* r, char* buf, size_t len);
/* Host-side lookup tables of a config, without a device (for checks):
 * disk: double[lut_size] = f^(1/4) (utils.rs:172-201);
 * blackbody: double[lut_size*3] linear-sRGB (blackbody.rs:56-124). */
kt_status kt_host_luts(const kt_config* cfg, double* disk_lut, double* bb_lut);

/* ---- lifecycle and frame API (src/renderer/mod.rs) ------------------- */
kt_status kt_create(const kt_config* cfg, kt_renderer** out);
void kt_destroy(kt_renderer* r);
kt_status kt_get_derived(const kt_renderer* r, kt_derived* out);
kt_status kt_submit_render(kt_renderer* r, const float cam_pos[3],
                           const float fwd[3], const float rgt[3],
                           const float up[3], float fov_scale,
                           int32_t* submitted);
kt_status kt_submit_render_f64(kt_renderer* r, const kt_camera* cam,
                               int32_t* submitted);
kt_status kt_ready_frame(kt_renderer* r, kt_frame_view* view,
                         int32_t* has_frame);
kt_status kt_finish_frame(kt_renderer* r, int32_t index);

/* ---- stream-ordered device API on caller-owned buffers --------------- */
/* hdr: float4[W*H] (rgb, 1); hit_class: uint8[W*H] or NULL.
 * Consecutive traces use four sets of launch state (work queue, counters,
 * tile order) in turn, so up to four traces issued on different streams
 * run concurrently on the device; a fifth waits for the first to finish. */
kt_status kt_trace(kt_renderer* r, const kt_camera* cam, float* hdr,
                   uint8_t* hit_class, void* stream);
/* As kt_trace, also writing each pixel's DP5(4) attempt count to
 * attempts (uint32[W*H], may be NULL).  New: the reference has no
 * per-ray work output. */
kt_status kt_trace_ex(kt_renderer* r, const kt_camera* cam, float* hdr,
                      uint8_t* hit_class, uint32_t* attempts, void* stream);
kt_status kt_set_launch_opts(kt_renderer* r, const kt_launch_opts* opts);
/* Waits for the last issued trace and returns its counters. */
kt_status kt_get_trace_stats(kt_renderer* r, kt_trace_stats* out);
/* bloom: float4[W*H] scratch (the passes use its first 12*W*H bytes as
 * packed RGB); ldr: uint32[W*H] 0x00RRGGBB. */
kt_status kt_post_process(kt_renderer* r, const float* hdr, float* bloom,
                          uint32_t* ldr, void* stream);
/* As kt_post_process, with the bloom intermediate in a device array the
 * library holds itself (24 MiB per device, kept until the process ends)
 * instead of a caller buffer.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops).  It calls no allocator.  Frames of up to 2,097,152
 * pixels (1920x1080 fits, 3840x2160 does not: KT_ERR_CAPACITY) when bloom
 * is on; any size when it is off.  All calls on one device share the
 * array: calls that overlap in time on one device (other streams, other
 * renderers, other threads) race and must be ordered by the caller. */
kt_status kt_post_process_noalloc(kt_renderer* r, const float* hdr,
                                  uint32_t* ldr, void* stream);
/* Host copies of the lookup tables the device uses (for parity checks):
 * disk: double[lut_size]; blackbody: double[lut_size*3] (rgb). */
kt_status kt_get_luts(const kt_renderer* r, double* disk_lut, double* bb_lut);

/* ---- batched persistent launch (new: the reference renders one frame per
 * launch) --------------------------------------------------------------
 * n renders with their own configs and cameras (spin, inclination,
 * resolution, ...; one precision mode for all) traced by one persistent
 * launch whose work queue holds the tiles of every render, hardest first.
 * hdr[i]: float4[W_i*H_i]; hit_class: NULL or an array of n pointers
 * (entries may be NULL).  Errors are read with kt_last_error(NULL, ...). */
typedef struct kt_batch kt_batch;
kt_status kt_batch_create(const kt_config* cfgs, const kt_camera* cams,
                          int32_t n, kt_batch** out);
void kt_batch_destroy(kt_batch* b);
kt_status kt_batch_set_launch_opts(kt_batch* b, const kt_launch_opts* opts);
kt_status kt_batch_trace(kt_batch* b, float* const* hdr,
                         uint8_t* const* hit_class, void* stream);
kt_status kt_batch_get_stats(kt_batch* b, kt_trace_stats* out);

/* ---- autotuner (new) ---------------------------------------------------
 * Traces the frame (kt_autotune: camera cam into the device buffer hdr;
 * kt_batch_autotune: every render into hdr[i]) with each of a fixed set
 * of launch options -- waves per CU, refill threshold (rays a wave keeps
 * before it takes new work), work-order tile edge -- keeps the fastest as
 * the handle's launch options and returns it in *best (may be NULL):
 * every candidate is timed twice (best run kept, HIP events around the
 * trace), then the three fastest four more times each, and the lowest
 * mean of those wins.  On an error the handle keeps its previous options.
 * The buffers hold the last trial's frame afterwards.  Synchronises the
 * stream. */
kt_status kt_autotune(kt_renderer* r, const kt_camera* cam, float* hdr,
                      void* stream, kt_launch_opts* best);
kt_status kt_batch_autotune(kt_batch* b, float* const* hdr, void* stream,
                            kt_launch_opts* best);

/* ---- differentiable renderer (new: the reference has no gradients) -----
 * The observer scene as a function of theta = {spin a, inclination i [rad],
 * observer distance r_obs [M], accretion-rate scale mdot}: camera at
 * r_obs (sin i, cos i, 0) looking at the hole (rgt = (0, 0, -1)), vertical
 * FOV opts->fov_deg, escape radius = r_obs, mass and every other knob from
 * cfg (exact right-hand side and spp = 1 are required; cfg->precision
 * chooses the forward tracer, FP64 or mixed).
 * kt_grad_forward traces the frame with the FP64 tracer and logs every
 * ray's committed steps (a state checkpoint every checkpoint_interval
 * steps); kt_grad_backward returns, for the frame of the last forward,
 *   grad[k] = sum over pixels and channels of dl_dhdr * d hdr / d theta_k
 * with the step sequence and every event (crossing, horizon, escape,
 * table cell) held fixed.  grad: device double[4]; grad_px: NULL or device
 * double[W*H*4] per-pixel contributions; dl_dhdr: device float4[W*H]
 * (alpha ignored).  Errors are read with kt_last_error(NULL, ...). */
enum {
    KT_ADJOINT_FP64 = 0,         /* every operation of the reverse pass in FP64 */
    KT_ADJOINT_FP32 = 1          /* step reversals (right-hand-side VJPs, stage
                                    and state adjoints) in FP32; re-integration,
                                    disk-crossing (dense-output root) and ray
                                    start derivatives in FP64 */
};
typedef struct kt_grad_opts {
    double fov_deg;              /* vertical field of view */
    int32_t checkpoint_interval; /* committed steps per checkpoint, 1..256 */
    int32_t log_steps_per_ray;   /* step-log room: W*H*(1 + ceil(this / interval)) chunks */
    int32_t adjoint_precision;   /* KT_ADJOINT_* */
    int32_t reserved;
} kt_grad_opts;
typedef struct kt_grad_stats {
    double forward_ms;           /* HIP events around the recording trace */
    double backward_ms;          /* HIP events around the adjoint pass and its reduction */
    uint64_t rhs_evals;          /* of the forward trace */
    uint64_t attempts;
    uint32_t chunks;             /* step-log chunks in use (W*H first chunks + pool) */
    uint32_t log_overflow;
    uint32_t checkpoint_interval;
    uint32_t reserved;
} kt_grad_stats;
typedef struct kt_grad kt_grad;
kt_status kt_grad_create(const kt_config* cfg, const kt_grad_opts* opts, kt_grad** out);
void kt_grad_destroy(kt_grad* g);
kt_status kt_grad_forward(kt_grad* g, const double theta[4], float* hdr,
                          uint8_t* hit_class, void* stream);
kt_status kt_grad_backward(kt_grad* g, const float* dl_dhdr, double* grad,
                           double* grad_px, void* stream);
kt_status kt_grad_get_stats(kt_grad* g, kt_grad_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* KERRTRACE_H */
