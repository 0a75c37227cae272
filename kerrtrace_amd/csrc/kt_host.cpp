// Host runtime of the MI355X renderer: configuration checks, the host-built
// lookup tables, device buffers, the double-buffered frame API and the
// C ABI declared in include/kerrtrace.h.
//
// Maps to the reference's Rust host:
//   derive()             <- build_kerr_params (src/renderer/utils.rs:53-114),
//                           validate_bloom_settings (src/renderer/init.rs:43-73)
//   disk_lut()           <- generate_disk_temperature_lut (utils.rs:172-201)
//   blackbody_lut()      <- generate_blackbody_lut (src/renderer/blackbody.rs:56-124)
//   kt_create            <- CudaRenderer::new (src/renderer/mod.rs:76-117)
//   kt_submit_render*    <- submit_render (mod.rs:119-218)
//   kt_ready_frame       <- ready_frame (mod.rs:220-250)
//   kt_finish_frame      <- finish_frame (mod.rs:252-260)
// The reference's textures (src/renderer/texture.rs) become plain device
// arrays read with explicit linear interpolation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <string>
#include <vector>

#include "../../include/kerrtrace.h"
#include "kt_adjoint.hpp"
#include "kt_host_physics.hpp"
#include "kt_internal.hpp"

namespace kt {
hipError_t launch_reset(DevCounters* c, uint32_t* error_flag, hipStream_t stream);
hipError_t launch_trace_fp64(const RenderParams* d_params, int n_renders, uint64_t total_items, int refill_threshold,
                             DevCounters* d_counters, int blocks, bool rhs_exact, hipStream_t stream);
hipError_t launch_trace_mixed(const RenderParams* d_params, int n_renders, uint64_t total_items,
                              int refill_threshold, DevCounters* d_counters, int blocks, bool rhs_exact,
                              hipStream_t stream);
hipError_t launch_bloom_h(const float* hdr, float* bloom, int W, int H, int R, const float* d_wts,
                          const float* d_ws_col, float threshold, hipStream_t stream);
hipError_t launch_post_v(const float* hdr, const float* bloom, uint32_t* out, int W, int H, int R, int K, int waves,
                         const float* d_wts, const float* d_ws_row, float intensity, int bloom_on,
                         hipStream_t stream);
bool post_v_plan(int R, size_t max_lds, int* k, int* waves);
hipError_t launch_cost_order(const RenderParams* d_params, int n_renders, int n_tiles, uint8_t* d_keys,
                             uint32_t* d_hist, uint32_t* d_perm, hipStream_t stream);
hipError_t launch_trace_fp64_record(const RenderParams* d_params, uint64_t total_items, int refill_threshold,
                                    DevCounters* d_counters, int blocks, hipStream_t stream);
hipError_t launch_reset_log(uint32_t* pool_head, uint32_t* overflow, uint32_t first_pool_chunk, hipStream_t stream);
hipError_t launch_grad_backward(const RenderParams* d_params, const GradConsts& G, const float* dl_dhdr,
                                double* grad_px, double* scratch, uint32_t threads, unsigned long long* queue,
                                const uint32_t* nsteps, uint8_t* keys, uint32_t* hist, uint32_t* order, int W, int H,
                                uint32_t n_pix, double* part, double* grad_out, bool mixed_fwd, bool fp32,
                                hipStream_t stream);
hipError_t launch_trace_mixed_record(const RenderParams* d_params, uint64_t total_items, int refill_threshold,
                                     DevCounters* d_counters, int blocks, hipStream_t stream);
size_t cost_order_hist_words(int n_tiles);
size_t bloom_h_lds_bytes(int R);
size_t noalloc_max_pixels();
hipError_t post_scratch_address(int device, float** out);
}  // namespace kt

// Mixed precision: a render whose tolerance is below this traces with the
// FP64 kernel.  FP32 rounds the right-hand side's gravitational terms to
// ~6e-8 relative; below ~5e-8 that rounding enters the embedded error
// estimate at the tolerance's own level, rejected steps shrink the step
// size towards the rounding floor and photon-ring rays take up to 4x more
// attempts (1920x1080, spin 0.998: 11712 against 2857 at 1e-9).
static constexpr double kMixedMinToleranceDefault = 5e-8;

namespace {

const double kPi = 3.14159265358979323846;
thread_local std::string g_last_error;

// ------------------------------------------------------------ host physics
using kt_host::calc_isco;
using kt_host::nt_factor;

// Multi-lobe Gaussian fit of the CIE 1931 colour-matching functions
// (Wyman, Sloan and Shirley 2013) as used by blackbody.rs:6-25.
double cmf_lobe(double x, double alpha, double mu, double s1, double s2) {
    double sigma = x < mu ? s1 : s2;
    double t = (x - mu) / sigma;
    return alpha * std::exp(-0.5 * t * t);
}

// Linear-sRGB colour of a blackbody for each LUT temperature
// (blackbody.rs:26-124): Planck radiance (c2 = 1.4388e7 nm K, scaled by
// 1e15, zero beyond exponent 80) times the CMFs, summed over the sampled
// wavelengths, then the IEC 61966-2-1 XYZ -> linear sRGB matrix.
void blackbody_lut(const kt_config& c, std::vector<double>& out) {
    const int n = (int)c.lut_size;
    out.assign((size_t)n * 3, 0.0);
    std::vector<double> lam, cx, cy, cz;
    for (double l = c.wavelength_start; !(l > c.wavelength_end); l += c.wavelength_step) {
        lam.push_back(l);
        cx.push_back(cmf_lobe(l, 1.056, 599.8, 37.9, 31.0) + cmf_lobe(l, 0.362, 442.0, 16.0, 26.7) +
                     cmf_lobe(l, -0.065, 501.1, 20.4, 26.2));
        cy.push_back(cmf_lobe(l, 0.821, 568.8, 46.9, 40.5) + cmf_lobe(l, 0.286, 530.9, 16.3, 31.1));
        cz.push_back(cmf_lobe(l, 1.217, 437.0, 11.8, 36.0) + cmf_lobe(l, 0.681, 459.0, 26.2, 13.8));
    }
    const double c2 = 1.4388e7;
    for (int i = 1; i < n; ++i) {
        double t = ((double)i / (double)(n - 1)) * c.lut_max_temp;
        if (t == 0.0) continue;
        double X = 0.0, Y = 0.0, Z = 0.0;
        for (size_t k = 0; k < lam.size(); ++k) {
            double l = lam[k];
            double e = c2 / (l * t);
            double rad = 0.0;
            if (!(e > 80.0)) {
                double l5 = l * l * l * l * l;
                rad = (1.0 / l5) / std::expm1(e) * 1e15;
            }
            X += rad * cx[k];
            Y += rad * cy[k];
            Z += rad * cz[k];
        }
        X *= c.wavelength_step;
        Y *= c.wavelength_step;
        Z *= c.wavelength_step;
        double* o = &out[(size_t)i * 3];
        o[0] = std::fma(3.240454, X, std::fma(-1.5371385, Y, -0.4985314 * Z));
        o[1] = std::fma(-0.969266, X, std::fma(1.8760108, Y, 0.041556 * Z));
        o[2] = std::fma(0.0556434, X, std::fma(-0.2040259, Y, 1.0572252 * Z));
    }
}

struct Derived {
    double rh, disk_inner, disk_outer, T_scale, lut_max;
    std::vector<double> disk_lut, bb_lut;
    int bloom_R;
    double bloom_sigma;
    bool bloom_active;
};

bool finite3(double a, double b, double c) { return std::isfinite(a) && std::isfinite(b) && std::isfinite(c); }

// Validates a config and derives everything the kernels read.
int derive(const kt_config& c, Derived& d, std::string& why) {
    if (c.width == 0 || c.height == 0) { why = "window width and height must be non-zero"; return KT_ERR_INVALID_CONFIG; }
    if (!std::isfinite(c.spin)) { why = "black hole spin is not finite"; return KT_ERR_INVALID_CONFIG; }
    if (!std::isfinite(c.mass)) { why = "black hole mass is not finite"; return KT_ERR_INVALID_CONFIG; }
    if (c.mass == 0.0) { why = "black hole mass must be non-zero"; return KT_ERR_INVALID_CONFIG; }
    if (!finite3(c.noise_scale, c.noise_strength, c.noise_winding)) { why = "disk noise parameter is not finite"; return KT_ERR_INVALID_CONFIG; }
    if (c.noise_detail > 0x7fffffffu) { why = "disk noise detail exceeds i32"; return KT_ERR_INVALID_CONFIG; }
    if (c.lut_size < 2) { why = "lut_size must be at least 2"; return KT_ERR_INVALID_CONFIG; }
    if (c.lut_size > (1u << 20)) { why = "lut_size above 2^20"; return KT_ERR_CAPACITY; }
    if (!(std::isfinite(c.wavelength_step) && c.wavelength_step > 0.0)) { why = "wavelength_step must be positive and finite"; return KT_ERR_INVALID_CONFIG; }
    if (!(std::isfinite(c.lut_max_temp) && c.lut_max_temp > 0.0)) { why = "lut_max_temp must be positive and finite"; return KT_ERR_INVALID_CONFIG; }
    if (!(std::isfinite(c.bloom_intensity) && c.bloom_intensity >= 0.0)) { why = "bloom intensity invalid"; return KT_ERR_INVALID_CONFIG; }
    if (!(std::isfinite(c.bloom_radius) && c.bloom_radius >= 0.0)) { why = "bloom radius invalid"; return KT_ERR_INVALID_CONFIG; }
    if (c.bloom_radius > 2147483647.0) { why = "bloom radius too large"; return KT_ERR_INVALID_CONFIG; }
    if (!(std::isfinite(c.bloom_threshold) && c.bloom_threshold >= 0.0)) { why = "bloom threshold invalid"; return KT_ERR_INVALID_CONFIG; }
    if (c.spp == 0) { why = "spp must be positive"; return KT_ERR_INVALID_CONFIG; }
    if (c.max_attempts == 0) { why = "max_attempts must be positive"; return KT_ERR_INVALID_CONFIG; }
    if (c.sky_grid_divisions == 0) { why = "sky grid divisions must be positive"; return KT_ERR_INVALID_CONFIG; }
    if (!(std::isfinite(c.mdot_scale) && c.mdot_scale >= 0.0)) { why = "mdot_scale must be finite and non-negative"; return KT_ERR_INVALID_CONFIG; }
    if (c.precision != KT_PRECISION_FP64 && c.precision != KT_PRECISION_MIXED) { why = "unknown precision mode"; return KT_ERR_INVALID_CONFIG; }
    if (c.disk_root_mode != KT_DISK_ROOT_SECANT_REF && c.disk_root_mode != KT_DISK_ROOT_TRUE) { why = "unknown disk root mode"; return KT_ERR_INVALID_CONFIG; }
    if (c.rhs_mode != KT_RHS_REFERENCE && c.rhs_mode != KT_RHS_HAMILTONIAN) { why = "unknown rhs mode"; return KT_ERR_INVALID_CONFIG; }
    if (!(c.mixed_min_tolerance >= 0.0)) { why = "mixed_min_tolerance must be >= 0"; return KT_ERR_INVALID_CONFIG; }
    if (c.precision == KT_PRECISION_MIXED && !(c.escape_radius <= 5000.0 * c.mass)) {
        why = "mixed precision needs escape_radius <= 5000 M (FP32 range of the right-hand side)";
        return KT_ERR_INVALID_CONFIG;
    }
    if ((uint64_t)c.width * c.height > (1ull << 31)) { why = "frame larger than 2^31 pixels"; return KT_ERR_CAPACITY; }
    const double M = c.mass;
    const double inv_m = 1.0 / M;
    const double aa = c.spin * c.spin;
    const double a_norm = c.spin * inv_m;
    d.rh = M + std::sqrt(std::fmax(std::fma(M, M, -aa), 0.0));
    d.disk_inner = calc_isco(a_norm, true, M);
    if (!std::isfinite(d.rh) || !std::isfinite(d.disk_inner)) { why = "Kerr parameters are not finite"; return KT_ERR_INVALID_CONFIG; }
    d.disk_outer = c.disk_outer_radius;
    if (!std::isfinite(d.disk_outer)) { why = "disk outer radius is not finite"; return KT_ERR_INVALID_CONFIG; }
    if (!(d.disk_outer > d.disk_inner)) {
        char b[160];
        std::snprintf(b, sizeof b, "disk outer radius must exceed the inner radius: outer=%g, inner=%g", d.disk_outer, d.disk_inner);
        why = b;
        return KT_ERR_INVALID_CONFIG;
    }
    const int n = (int)c.lut_size;
    d.disk_lut.resize(n);
    const double span = d.disk_outer - d.disk_inner;
    d.lut_max = 0.0;
    for (int i = 0; i < n; ++i) {
        double r = std::fma((double)i / (double)(n - 1), span, d.disk_inner);
        double f = nt_factor(r, a_norm, d.disk_inner, inv_m);
        if (!std::isfinite(f)) { why = "disk flux factor is not finite"; return KT_ERR_INVALID_CONFIG; }
        double v = std::sqrt(std::sqrt(f));
        d.disk_lut[i] = v;
        d.lut_max = std::max(d.lut_max, v);
    }
    double T_scale = c.disk_temperature_scale;
    if (c.disk_peak_temperature > 0.0) {
        if (!(d.lut_max > 0.0)) { why = "disk flux LUT is identically zero"; return KT_ERR_INVALID_CONFIG; }
        T_scale = c.disk_peak_temperature / d.lut_max;
    }
    d.T_scale = T_scale * std::sqrt(std::sqrt(c.mdot_scale));
    blackbody_lut(c, d.bb_lut);
    d.bloom_R = c.bloom_radius > 0.0 ? (int)std::ceil(c.bloom_radius) : 0;
    d.bloom_sigma = c.bloom_radius / 3.0;
    d.bloom_active = c.bloom_enabled != 0 && d.bloom_R > 0 && c.bloom_intensity > 0.0;
    return KT_OK;
}

void fill_derived(const Derived& d, kt_derived* out) {
    out->rh = d.rh;
    out->disk_inner = d.disk_inner;
    out->disk_outer = d.disk_outer;
    out->temperature_scale = d.T_scale;
    out->disk_lut_max = d.lut_max;
    out->bloom_sigma = d.bloom_sigma;
    out->bloom_radius_int = d.bloom_R;
    out->bloom_active = d.bloom_active ? 1 : 0;
}

}  // namespace

// --------------------------------------------------------------- renderer
struct kt_renderer {
    kt_config cfg;
    Derived d;
    std::string err;
    int device = 0;
    int cus = 256;
    size_t max_lds = 65536;
    double* d_disk_lut = nullptr;
    double* d_bb_lut = nullptr;
    float* d_disk_lut_f = nullptr;
    float* d_bb_lut_f = nullptr;
    float* d_bloom_w = nullptr;
    float* d_bloom_ws = nullptr;  // [W] column then [H] row sums of in-bounds weights
    float* d_noalloc_bloom = nullptr;  // the library's static bloom array on this device (not owned)
    std::string noalloc_err;           // why d_noalloc_bloom could not be resolved
    int post_k = 8, post_waves = 8;  // post_v tile: rows per lane, waves per block
    // Per-launch state in kSlots slots that traces use in turn, so that up
    // to kSlots traces issued on as many streams run concurrently: the
    // long-ray tail of one frame then overlaps the bulk of the next.  A slot
    // is reused only after its previous trace has finished (stream wait on
    // `done`).
    struct Slot {
        kt::DevCounters* counters = nullptr;
        uint32_t* flag = nullptr;       // LUT-overflow flag of kt_trace launches
        uint32_t* flag_used = nullptr;  // the flag the slot's last trace wrote
        // hardest-first tile order: keys, bucket counts/offsets, permutation
        uint8_t* keys = nullptr;
        uint32_t* hist = nullptr;
        uint32_t* perm = nullptr;
        hipEvent_t ev0 = nullptr, done = nullptr;  // around the tracer launch
        bool used = false;
        uint64_t rays = 0;
    };
    static constexpr int kSlots = 4;
    Slot slots[kSlots];
    int next_slot = 0, last_slot = -1;
    size_t max_tiles = 0;
    // ring of parameter records: pinned host staging -> device
    static constexpr int kRing = 16;
    kt::RenderParams* h_params = nullptr;
    kt::RenderParams* d_params = nullptr;
    hipEvent_t param_ev[kRing] = {};
    bool param_ev_used[kRing] = {};
    int param_next = 0;
    kt_launch_opts opts{8, 8, 8, 1, 256, 0};
    // double-buffered frames (mod.rs:24-51, 88-91), each on its own stream
    struct Frame {
        float* hdr = nullptr;
        float* bloom = nullptr;
        uint32_t* ldr = nullptr;
        uint32_t* ldr_host = nullptr;
        uint32_t* flag = nullptr;
        uint32_t* flag_host = nullptr;
        hipEvent_t ready = nullptr;
        hipStream_t stream = nullptr;
        bool pending = false;
    };
    Frame frames[2];
    std::deque<int> queue;
    int next_frame = 0;
};

namespace {

// Whether a render traces with the mixed kernel (FP32 right-hand-side bulk):
// mixed precision and a tolerance at or above its FP32 floor.
bool uses_fp32_rhs(const kt_config& c) {
    return c.precision == KT_PRECISION_MIXED && c.tolerance >= c.mixed_min_tolerance;
}

int set_err(kt_renderer* r, int code, const std::string& msg) {
    if (r) r->err = msg;
    g_last_error = msg;
    return code;
}

int hip_err(kt_renderer* r, hipError_t e, const char* what) {
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return set_err(r, KT_ERR_DEVICE, m);
}

#define KT_HIP(r, call, what)                      \
    do {                                           \
        hipError_t e_ = (call);                    \
        if (e_ != hipSuccess) return hip_err(r, e_, what); \
    } while (0)

void release(kt_renderer* r) {
    if (!r) return;
    for (auto& f : r->frames) {
        if (f.hdr) (void)hipFree(f.hdr);
        if (f.bloom) (void)hipFree(f.bloom);
        if (f.ldr) (void)hipFree(f.ldr);
        if (f.flag) (void)hipFree(f.flag);
        if (f.ldr_host) (void)hipHostFree(f.ldr_host);
        if (f.flag_host) (void)hipHostFree(f.flag_host);
        if (f.ready) (void)hipEventDestroy(f.ready);
        if (f.stream) (void)hipStreamDestroy(f.stream);
    }
    for (auto& sl : r->slots) {
        if (sl.counters) (void)hipFree(sl.counters);
        if (sl.flag) (void)hipFree(sl.flag);
        if (sl.keys) (void)hipFree(sl.keys);
        if (sl.hist) (void)hipFree(sl.hist);
        if (sl.perm) (void)hipFree(sl.perm);
        if (sl.ev0) (void)hipEventDestroy(sl.ev0);
        if (sl.done) (void)hipEventDestroy(sl.done);
    }
    for (int i = 0; i < kt_renderer::kRing; ++i)
        if (r->param_ev[i]) (void)hipEventDestroy(r->param_ev[i]);
    if (r->h_params) (void)hipHostFree(r->h_params);
    if (r->d_params) (void)hipFree(r->d_params);
    if (r->d_disk_lut) (void)hipFree(r->d_disk_lut);
    if (r->d_bb_lut) (void)hipFree(r->d_bb_lut);
    if (r->d_disk_lut_f) (void)hipFree(r->d_disk_lut_f);
    if (r->d_bb_lut_f) (void)hipFree(r->d_bb_lut_f);
    if (r->d_bloom_w) (void)hipFree(r->d_bloom_w);
    if (r->d_bloom_ws) (void)hipFree(r->d_bloom_ws);
    delete r;
}

// The parameter record of one render with camera `cam` writing to hdr/cls.
void make_params(const kt_renderer* r, const kt_config& c, const Derived& d, const kt_camera& cam, float* hdr,
                 uint8_t* cls, uint32_t* flag, int tile, kt::RenderParams& P, uint32_t* attempts = nullptr) {
    std::memset(&P, 0, sizeof P);
    P.a = c.spin;
    P.M = c.mass;
    P.aa = c.spin * c.spin;
    P.inv_M = 1.0 / c.mass;
    P.rh_eps = d.rh + c.horizon_epsilon;
    P.disk_inner = d.disk_inner;
    P.disk_outer = d.disk_outer;
    P.T_scale = d.T_scale;
    P.h0 = c.initial_step;
    P.tol = c.tolerance;
    P.tol_floor = c.tol_floor;
    P.cutoff = c.transmittance_cutoff;
    P.R_esc = c.escape_radius;
    P.inv_w_2 = 2.0 / (double)c.width;
    P.inv_h_2 = 2.0 / (double)c.height;
    P.aspect = (double)c.width / (double)c.height;
    P.exposure_over_spp = c.exposure_scale * (1.0 / (double)c.spp);
    for (int k = 0; k < 3; ++k) {
        P.cam[k] = cam.pos[k];
        P.fwd[k] = cam.fwd[k];
        P.rgt[k] = cam.rgt[k];
        P.up[k] = cam.up[k];
    }
    P.fov_scale = cam.fov_scale;
    P.sky_spacing = kPi / (double)c.sky_grid_divisions;
    P.sky_lt = c.sky_line_thickness;
    P.sky_int = c.sky_intensity;
    P.noise_scale = c.noise_scale;
    P.noise_strength = c.noise_strength;
    P.noise_winding = c.noise_winding;
    P.noise_seed_mix = c.noise_seed * 0x9E3779B97F4A7C15ULL;
    P.wl_step = c.wavelength_step;
    P.lut_max_temp = c.lut_max_temp;
    P.disk_lut = r ? r->d_disk_lut : nullptr;
    P.bb_lut = r ? r->d_bb_lut : nullptr;
    P.disk_lut_f = r ? r->d_disk_lut_f : nullptr;
    P.bb_lut_f = r ? r->d_bb_lut_f : nullptr;
    P.hdr = hdr;
    P.cls = cls;
    P.attempts_out = attempts;
    P.error_flag = flag;
    P.W = (int)c.width;
    P.H = (int)c.height;
    P.spp = (int)c.spp;
    P.max_steps = (int)std::min<uint32_t>(c.max_steps, 0x7fffffffu);
    P.max_attempts = (int)std::min<uint32_t>(c.max_attempts, 0x7fffffffu);
    P.lut_n = (int)c.lut_size;
    bool noise_on = c.noise_enabled != 0 && c.noise_detail > 0 && c.noise_strength != 0.0;
    P.noise_detail = noise_on ? (int)c.noise_detail : 0;
    P.disk_root_mode = c.disk_root_mode;
    P.rhs_exact = c.rhs_mode == KT_RHS_HAMILTONIAN ? 1 : 0;
    P.tile = tile;
    P.tiles_x = ((int)c.width + tile - 1) / tile;
    int tiles_y = ((int)c.height + tile - 1) / tile;
    P.tile_begin = 0;
    P.n_tiles = (uint32_t)(P.tiles_x * tiles_y);
}

// Stages a parameter record through the pinned ring and returns its
// device address (stream-ordered).  The slot's event is recorded after the
// copy here and again by the caller after the kernels that read the record;
// a slot is reused only once that event has completed.
int upload_params(kt_renderer* r, const kt::RenderParams& P, hipStream_t s, kt::RenderParams** out) {
    int i = r->param_next;
    r->param_next = (i + 1) % kt_renderer::kRing;
    if (r->param_ev_used[i]) KT_HIP(r, hipEventSynchronize(r->param_ev[i]), "param ring wait");
    r->h_params[i] = P;
    KT_HIP(r, hipMemcpyAsync(&r->d_params[i], &r->h_params[i], sizeof P, hipMemcpyHostToDevice, s), "param upload");
    KT_HIP(r, hipEventRecord(r->param_ev[i], s), "param event");
    r->param_ev_used[i] = true;
    *out = &r->d_params[i];
    return KT_OK;
}

// Traces one frame on stream s with the next per-launch slot; flag null
// means the slot's own LUT-overflow flag.
int trace_into(kt_renderer* r, const kt_camera* cam, float* hdr, uint8_t* cls, uint32_t* flag, hipStream_t s,
               uint32_t* attempts = nullptr) {
    kt_renderer::Slot& sl = r->slots[r->next_slot];
    if (!flag) flag = sl.flag;
    if (sl.used) KT_HIP(r, hipStreamWaitEvent(s, sl.done, 0), "slot wait");
    kt::RenderParams P;
    int tile = r->opts.queue_chunk > 0 ? r->opts.queue_chunk : 8;
    make_params(r, r->cfg, r->d, *cam, hdr, cls, flag, tile, P, attempts);
    P.long_ray = r->opts.long_ray_attempts > 0 ? (uint32_t)r->opts.long_ray_attempts : 256u;
    P.isolate = (uint32_t)r->opts.isolate_level;
    const bool order = r->opts.cost_order != 2 && tile <= 8;
    const int n_tiles = (int)P.n_tiles;
    const uint64_t items = (uint64_t)n_tiles * (uint64_t)tile * (uint64_t)tile;
    if (order) P.tile_perm = sl.perm;
    kt::RenderParams* dp = nullptr;
    int rc = upload_params(r, P, s, &dp);
    if (rc != KT_OK) return rc;
    if (order) {
        hipError_t e = kt::launch_cost_order(dp, 1, n_tiles, sl.keys, sl.hist, sl.perm, s);
        if (e != hipSuccess) return set_err(r, KT_ERR_LAUNCH, std::string("cost order launch: ") + hipGetErrorString(e));
    }
    KT_HIP(r, kt::launch_reset(sl.counters, flag, s), "counter reset launch");
    int wpc = r->opts.waves_per_cu > 0 ? r->opts.waves_per_cu : 8;
    int blocks = std::max(1, (r->cus * wpc + 3) / 4);
    int thr = r->opts.refill_threshold > 0 ? std::min(r->opts.refill_threshold, 64) : 8;
    KT_HIP(r, hipEventRecord(sl.ev0, s), "trace event");
    hipError_t e = uses_fp32_rhs(r->cfg)
                       ? kt::launch_trace_mixed(dp, 1, items, thr, sl.counters, blocks, P.rhs_exact != 0, s)
                       : kt::launch_trace_fp64(dp, 1, items, thr, sl.counters, blocks, P.rhs_exact != 0, s);
    if (e != hipSuccess) return set_err(r, KT_ERR_LAUNCH, std::string("trace launch: ") + hipGetErrorString(e));
    KT_HIP(r, hipEventRecord(sl.done, s), "trace event");
    // the ring slot's event now follows the kernels that read the record, so
    // the record is not overwritten under a trace still running on another
    // stream when the ring comes round
    KT_HIP(r, hipEventRecord(r->param_ev[dp - r->d_params], s), "param event");
    sl.used = true;
    sl.flag_used = flag;
    sl.rays = (uint64_t)r->cfg.width * r->cfg.height * r->cfg.spp;
    r->last_slot = r->next_slot;
    r->next_slot = (r->next_slot + 1) % kt_renderer::kSlots;
    return KT_OK;
}

int post_into(kt_renderer* r, const float* hdr, float* bloom, uint32_t* ldr, hipStream_t s) {
    const kt_config& c = r->cfg;
    int W = (int)c.width, H = (int)c.height, R = r->d.bloom_R;
    bool on = r->d.bloom_active;
    if (on) {
        hipError_t e = kt::launch_bloom_h(hdr, bloom, W, H, R, r->d_bloom_w, r->d_bloom_ws,
                                          (float)c.bloom_threshold, s);
        if (e != hipSuccess) return set_err(r, KT_ERR_LAUNCH, std::string("bloom launch: ") + hipGetErrorString(e));
    }
    hipError_t e = kt::launch_post_v(hdr, bloom, ldr, W, H, R, r->post_k, r->post_waves, r->d_bloom_w,
                                     r->d_bloom_ws ? r->d_bloom_ws + W : nullptr, (float)c.bloom_intensity,
                                     on ? 1 : 0, s);
    if (e != hipSuccess) return set_err(r, KT_ERR_LAUNCH, std::string("post launch: ") + hipGetErrorString(e));
    return KT_OK;
}

}  // namespace

// ------------------------------------------------------------ batch launch
// Many renders (different spins, cameras, resolutions) traced by one
// persistent launch: one parameter record per render, their tiles numbered
// consecutively in one work queue (hardest-first across all renders).
struct kt_batch {
    int n = 0;
    std::vector<kt_config> cfgs;
    std::vector<kt_camera> cams;
    std::vector<Derived> ds;
    std::string err;
    int device = 0;
    int cus = 256;
    int precision = KT_PRECISION_FP64;
    std::vector<double*> d_disk_luts;
    std::vector<double*> d_bb_luts;
    kt::RenderParams* h_params = nullptr;  // pinned
    kt::RenderParams* d_params = nullptr;
    kt::DevCounters* d_counters = nullptr;
    uint32_t* d_flag = nullptr;
    uint8_t* d_keys = nullptr;
    uint32_t* d_hist = nullptr;
    uint32_t* d_perm = nullptr;
    size_t max_tiles = 0;
    kt_launch_opts opts{8, 8, 8, 1, 256, 0};
    hipStream_t last_stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_params = nullptr;
    bool params_in_flight = false;
    uint64_t last_rays = 0;
};

namespace {

void release_batch(kt_batch* b) {
    if (!b) return;
    for (double* p : b->d_disk_luts) if (p) (void)hipFree(p);
    for (double* p : b->d_bb_luts) if (p) (void)hipFree(p);
    if (b->h_params) (void)hipHostFree(b->h_params);
    if (b->d_params) (void)hipFree(b->d_params);
    if (b->d_counters) (void)hipFree(b->d_counters);
    if (b->d_flag) (void)hipFree(b->d_flag);
    if (b->d_keys) (void)hipFree(b->d_keys);
    if (b->d_hist) (void)hipFree(b->d_hist);
    if (b->d_perm) (void)hipFree(b->d_perm);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->ev_params) (void)hipEventDestroy(b->ev_params);
    delete b;
}

bool same_blackbody(const kt_config& a, const kt_config& b) {
    return a.lut_size == b.lut_size && a.lut_max_temp == b.lut_max_temp && a.wavelength_start == b.wavelength_start &&
           a.wavelength_end == b.wavelength_end && a.wavelength_step == b.wavelength_step;
}

}  // namespace

// ------------------------------------------------ differentiable renderer
// The observer scene as a function of theta = (spin, inclination [rad],
// r_obs, mdot): one recording FP64 trace per forward, one adjoint pass per
// backward (kt_grad.hip).
struct kt_grad {
    kt_config cfg;
    kt_grad_opts opts;
    Derived d;
    double theta[4] = {0, 0, 0, 0};
    double d_rin_da = 0, dTs_da = 0, dTs_dm = 0;
    std::vector<double> dlut;
    int device = 0, cus = 256;
    uint32_t n_pix = 0, K = 32, capacity = 0, stride = 0;
    uint32_t threads = 0;  // adjoint grid
    double* d_disk_lut = nullptr;
    double* d_bb_lut = nullptr;
    double* d_dlut = nullptr;
    kt::RenderParams* h_params = nullptr;
    kt::RenderParams* d_params = nullptr;
    kt::DevCounters* d_counters = nullptr;
    uint32_t* d_flag = nullptr;
    uint8_t* d_cls = nullptr;
    uint8_t* d_keys = nullptr;
    uint32_t* d_hist = nullptr;
    uint32_t* d_perm = nullptr;
    double* d_arena = nullptr;
    uint32_t* d_head = nullptr;
    uint32_t* d_nsteps = nullptr;
    double* d_pk = nullptr;
    uint32_t* d_pool = nullptr;  // [0] pool head, [1] overflow
    uint32_t* h_pool = nullptr;  // pinned copy
    double* d_grad_px = nullptr;
    double* d_scratch = nullptr;
    uint8_t* d_rkeys = nullptr;   // adjoint order: per-pixel step-count keys,
    uint32_t* d_rhist = nullptr;  // sort counters,
    uint32_t* d_rorder = nullptr; // tiles by descending longest ray
    double* d_part = nullptr;     // partial gradient sums
    unsigned long long* d_queue = nullptr;
    hipEvent_t ev[4] = {};
    hipEvent_t ev_params = nullptr;
    bool params_pending = false, have_forward = false, have_backward = false;
    hipStream_t last_stream = nullptr;  // stream of the last forward or backward
    hipStream_t fwd_stream = nullptr;   // stream of the last forward (its class copy runs there)
    uint64_t items = 0;
};

namespace {

void release_grad(kt_grad* g) {
    if (!g) return;
    void* bufs[] = {g->d_disk_lut, g->d_bb_lut, g->d_dlut, g->d_params, g->d_counters, g->d_flag, g->d_cls,
                    g->d_keys, g->d_hist, g->d_perm, g->d_arena, g->d_head, g->d_nsteps, g->d_pk, g->d_pool,
                    g->d_grad_px, g->d_scratch, g->d_queue, g->d_rkeys, g->d_rhist, g->d_rorder, g->d_part};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    if (g->h_params) (void)hipHostFree(g->h_params);
    if (g->h_pool) (void)hipHostFree(g->h_pool);
    for (hipEvent_t e : g->ev)
        if (e) (void)hipEventDestroy(e);
    if (g->ev_params) (void)hipEventDestroy(g->ev_params);
    delete g;
}

// Values and spin / mdot derivatives of everything the scene derives from
// theta on the host: disk_inner, the disk LUT, temperature_scale.
int derive_grad(kt_grad* g, const double theta[4], std::string& why) {
    kt_config c = g->cfg;
    c.spin = theta[0];
    c.mdot_scale = theta[3];
    c.escape_radius = theta[2];
    int rc = derive(c, g->d, why);
    if (rc != KT_OK) return rc;
    using kt_host::HD;
    const double M = c.mass, inv_m = 1.0 / M;
    const HD a_norm{c.spin * inv_m, inv_m};
    const HD rin = kt_host::calc_isco(a_norm, true, M);
    g->d_rin_da = rin.d;
    const int n = (int)c.lut_size;
    g->dlut.assign(n, 0.0);
    const HD span = HD{c.disk_outer_radius, 0.0} - rin;
    int imax = 0;
    for (int i = 0; i < n; ++i) {
        HD r = kt_host::m_fma(HD{(double)i / (double)(n - 1), 0.0}, span, rin);
        HD f = kt_host::nt_factor(r, a_norm, rin, inv_m);
        if (f.v > 0.0) {
            HD v = kt_host::m_sqrt(kt_host::m_sqrt(f));
            g->dlut[i] = v.d;
        }
        if (g->d.disk_lut[i] > g->d.disk_lut[imax]) imax = i;
    }
    const double Ts = g->d.T_scale;
    g->dTs_da = (c.disk_peak_temperature > 0.0 && g->d.lut_max > 0.0) ? -Ts * g->dlut[imax] / g->d.lut_max : 0.0;
    g->dTs_dm = c.mdot_scale > 0.0 ? Ts / (4.0 * c.mdot_scale) : 0.0;
    for (int k = 0; k < 4; ++k) g->theta[k] = theta[k];
    return KT_OK;
}

}  // namespace

// ==================================================================== C ABI
extern "C" {

kt_status kt_config_defaults(kt_config* c) {
    if (!c) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null config");
    std::memset(c, 0, sizeof *c);
    c->width = 2100;
    c->height = 900;
    c->spp = 4;
    c->sky_grid_divisions = 40;
    c->exposure_scale = 3e-3;
    c->sky_line_thickness = 5e-4;
    c->sky_intensity = 30.0;
    c->spin = -0.45;
    c->mass = 0.5;
    c->disk_outer_radius = 40.0;
    c->disk_temperature_scale = 3e4;
    c->initial_step = 1.0;
    c->tolerance = 1e-5;
    c->max_steps = 100;
    c->max_attempts = 3;
    c->transmittance_cutoff = 1e-2;
    c->horizon_epsilon = 5e-7;
    c->escape_radius = 80.0;
    c->lut_size = 4096;
    c->lut_max_temp = 3e4;
    c->wavelength_start = 380.0;
    c->wavelength_end = 720.0;
    c->wavelength_step = 10.0;
    c->noise_enabled = 1;
    c->noise_detail = 10;
    c->noise_scale = 0.1;
    c->noise_strength = 0.6;
    c->noise_winding = 50.0;
    c->noise_seed = 0;
    c->bloom_enabled = 1;
    c->bloom_intensity = 0.5;
    c->bloom_radius = 48.0;
    c->bloom_threshold = 0.0;
    c->tol_floor = 0.0;
    c->disk_peak_temperature = 0.0;
    c->mdot_scale = 1.0;
    c->disk_root_mode = KT_DISK_ROOT_SECANT_REF;
    c->precision = KT_PRECISION_FP64;
    c->rhs_mode = KT_RHS_REFERENCE;
    c->mixed_min_tolerance = kMixedMinToleranceDefault;
    return KT_OK;
}

kt_status kt_validate_config(const kt_config* c, kt_derived* out) {
    if (!c) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null config");
    Derived d;
    std::string why;
    int rc = derive(*c, d, why);
    if (rc != KT_OK) return set_err(nullptr, rc, why);
    if (out) fill_derived(d, out);
    return KT_OK;
}

kt_status kt_host_luts(const kt_config* c, double* disk, double* bb) {
    if (!c) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null config");
    Derived d;
    std::string why;
    int rc = derive(*c, d, why);
    if (rc != KT_OK) return set_err(nullptr, rc, why);
    if (disk) std::memcpy(disk, d.disk_lut.data(), d.disk_lut.size() * sizeof(double));
    if (bb) std::memcpy(bb, d.bb_lut.data(), d.bb_lut.size() * sizeof(double));
    return KT_OK;
}

kt_status kt_camera_from_yaw_pitch(const double pos[3], double yaw_deg, double pitch_deg, double fov_deg,
                                   kt_camera* out) {
    if (!pos || !out) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    double yaw = yaw_deg * (kPi / 180.0), pitch = pitch_deg * (kPi / 180.0);
    double f[3] = {std::sin(yaw) * std::cos(pitch), std::sin(pitch), -std::cos(yaw) * std::cos(pitch)};
    auto norm = [](double v[3]) {
        double n = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        v[0] /= n; v[1] /= n; v[2] /= n;
    };
    norm(f);
    double rgt[3] = {f[1] * 0.0 - f[2] * 1.0, f[2] * 0.0 - f[0] * 0.0, f[0] * 1.0 - f[1] * 0.0};
    norm(rgt);
    double up[3] = {rgt[1] * f[2] - rgt[2] * f[1], rgt[2] * f[0] - rgt[0] * f[2], rgt[0] * f[1] - rgt[1] * f[0]};
    norm(up);
    for (int k = 0; k < 3; ++k) {
        out->pos[k] = pos[k];
        out->fwd[k] = f[k];
        out->rgt[k] = rgt[k];
        out->up[k] = up[k];
    }
    out->fov_scale = std::tan(fov_deg * (kPi / 180.0) / 2.0);
    return KT_OK;
}

kt_status kt_camera_observer(double r_obs, double incl_deg, double fov_deg, kt_camera* out) {
    if (!out) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null camera");
    if (!(std::isfinite(r_obs) && r_obs > 0.0 && std::isfinite(incl_deg) && std::isfinite(fov_deg)))
        return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "observer parameters must be finite, r_obs > 0");
    double i = incl_deg * (kPi / 180.0);
    double si = std::sin(i), ci = std::cos(i);
    double pos[3] = {r_obs * si, r_obs * ci, 0.0};
    double fwd[3] = {-si, -ci, 0.0};
    double rgt[3] = {0.0, 0.0, -1.0};
    double up[3] = {-ci, si, 0.0};
    for (int k = 0; k < 3; ++k) {
        out->pos[k] = pos[k];
        out->fwd[k] = fwd[k];
        out->rgt[k] = rgt[k];
        out->up[k] = up[k];
    }
    out->fov_scale = std::tan(fov_deg * (kPi / 180.0) / 2.0);
    return KT_OK;
}

kt_status kt_last_error(const kt_renderer* r, char* buf, size_t len) {
    if (!buf || len == 0) return KT_ERR_INVALID_ARGUMENT;
    const std::string& m = (r && !r->err.empty()) ? r->err : g_last_error;
    std::snprintf(buf, len, "%s", m.c_str());
    return KT_OK;
}

kt_status kt_create(const kt_config* cfg, kt_renderer** out) {
    if (!cfg || !out) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    kt_renderer* r = new (std::nothrow) kt_renderer();
    if (!r) return set_err(nullptr, KT_ERR_DEVICE, "out of host memory");
    r->cfg = *cfg;
    std::string why;
    int rc = derive(r->cfg, r->d, why);
    if (rc != KT_OK) {
        release(r);
        return set_err(nullptr, rc, why);
    }
#define KT_CREATE_HIP(call, what)                 \
    do {                                          \
        hipError_t e_ = (call);                   \
        if (e_ != hipSuccess) {                   \
            int rc_ = hip_err(nullptr, e_, what); \
            release(r);                           \
            return rc_;                           \
        }                                         \
    } while (0)
    KT_CREATE_HIP(hipGetDevice(&r->device), "hipGetDevice");
    hipDeviceProp_t prop;
    KT_CREATE_HIP(hipGetDeviceProperties(&prop, r->device), "hipGetDeviceProperties");
    r->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    r->max_lds = prop.sharedMemPerBlock > 0 ? prop.sharedMemPerBlock : 65536;
    {
        // device address of the static bloom array of kt_post_process_noalloc
        // (resolved once per device; a failure only disables that entry point)
        hipError_t e = kt::post_scratch_address(r->device, &r->d_noalloc_bloom);
        if (e != hipSuccess) {
            r->d_noalloc_bloom = nullptr;
            r->noalloc_err = std::string("hipGetSymbolAddress: ") + hipGetErrorString(e);
            (void)hipGetLastError();
        }
    }
    const int n = (int)cfg->lut_size;
    std::vector<float> disk_f(n), bb_f((size_t)n * 3);
    for (int i = 0; i < n; ++i) disk_f[i] = (float)r->d.disk_lut[i];
    for (size_t i = 0; i < bb_f.size(); ++i) bb_f[i] = (float)r->d.bb_lut[i];
    KT_CREATE_HIP(hipMalloc(&r->d_disk_lut, sizeof(double) * n), "hipMalloc disk LUT");
    KT_CREATE_HIP(hipMalloc(&r->d_bb_lut, sizeof(double) * n * 3), "hipMalloc blackbody LUT");
    KT_CREATE_HIP(hipMalloc(&r->d_disk_lut_f, sizeof(float) * n), "hipMalloc disk LUT f32");
    KT_CREATE_HIP(hipMalloc(&r->d_bb_lut_f, sizeof(float) * n * 3), "hipMalloc blackbody LUT f32");
    KT_CREATE_HIP(hipMemcpy(r->d_disk_lut, r->d.disk_lut.data(), sizeof(double) * n, hipMemcpyHostToDevice), "upload disk LUT");
    KT_CREATE_HIP(hipMemcpy(r->d_bb_lut, r->d.bb_lut.data(), sizeof(double) * n * 3, hipMemcpyHostToDevice), "upload blackbody LUT");
    KT_CREATE_HIP(hipMemcpy(r->d_disk_lut_f, disk_f.data(), sizeof(float) * n, hipMemcpyHostToDevice), "upload disk LUT f32");
    KT_CREATE_HIP(hipMemcpy(r->d_bb_lut_f, bb_f.data(), sizeof(float) * n * 3, hipMemcpyHostToDevice), "upload blackbody LUT f32");
    // bloom tap weights exp(-i^2 / (2 sigma^2)) (kernel.cu:48-53)
    const int R = r->d.bloom_R;
    if (r->d.bloom_active) {
        if (kt::bloom_h_lds_bytes(R) > r->max_lds) {
            release(r);
            return set_err(nullptr, KT_ERR_CAPACITY, "bloom radius too large for the LDS-staged blur");
        }
        if (!kt::post_v_plan(R, r->max_lds, &r->post_k, &r->post_waves)) {
            release(r);
            return set_err(nullptr, KT_ERR_CAPACITY, "bloom radius too large for the LDS-staged blur");
        }
    }
    std::vector<float> w((size_t)2 * R + 1, 0.f);
    for (int i = -R; i <= R; ++i) {
        double t = (double)i;
        w[i + R] = (float)std::exp(-(t * t) / (2.0 * r->d.bloom_sigma * r->d.bloom_sigma));
    }
    KT_CREATE_HIP(hipMalloc(&r->d_bloom_w, sizeof(float) * w.size()), "hipMalloc bloom weights");
    KT_CREATE_HIP(hipMemcpy(r->d_bloom_w, w.data(), sizeof(float) * w.size(), hipMemcpyHostToDevice), "upload bloom weights");
    if (r->d.bloom_active) {
        // renormalisation: the sum of the in-bounds tap weights of each
        // column and row, accumulated in tap order (kernel.cu:54-84)
        const int W = (int)cfg->width, H = (int)cfg->height;
        std::vector<float> ws((size_t)W + H);
        auto in_bounds_sum = [&](int p, int n) {
            float acc = 0.f;
            for (int i = std::max(-R, -p); i <= std::min(R, n - 1 - p); ++i) acc += w[i + R];
            return acc;
        };
        for (int x = 0; x < W; ++x) ws[x] = in_bounds_sum(x, W);
        for (int y = 0; y < H; ++y) ws[(size_t)W + y] = in_bounds_sum(y, H);
        KT_CREATE_HIP(hipMalloc(&r->d_bloom_ws, sizeof(float) * ws.size()), "hipMalloc bloom weight sums");
        KT_CREATE_HIP(hipMemcpy(r->d_bloom_ws, ws.data(), sizeof(float) * ws.size(), hipMemcpyHostToDevice),
                      "upload bloom weight sums");
    }
    r->max_tiles = (size_t)cfg->width * cfg->height;  // tile edge 1 is the largest count
    for (auto& sl : r->slots) {
        KT_CREATE_HIP(hipMalloc(&sl.counters, sizeof(kt::DevCounters)), "hipMalloc counters");
        KT_CREATE_HIP(hipMemset(sl.counters, 0, sizeof(kt::DevCounters)), "clear counters");
        KT_CREATE_HIP(hipMalloc(&sl.flag, 2 * sizeof(uint32_t)), "hipMalloc flag");
        KT_CREATE_HIP(hipMemset(sl.flag, 0, 2 * sizeof(uint32_t)), "clear flag");
        KT_CREATE_HIP(hipMalloc(&sl.keys, r->max_tiles), "hipMalloc cost keys");
        KT_CREATE_HIP(hipMalloc(&sl.hist, kt::cost_order_hist_words((int)r->max_tiles) * sizeof(uint32_t)),
                      "hipMalloc cost histogram");
        KT_CREATE_HIP(hipMalloc(&sl.perm, r->max_tiles * sizeof(uint32_t)), "hipMalloc tile order");
        KT_CREATE_HIP(hipEventCreate(&sl.ev0), "trace event");
        KT_CREATE_HIP(hipEventCreate(&sl.done), "trace event");
    }
    KT_CREATE_HIP(hipHostMalloc(&r->h_params, sizeof(kt::RenderParams) * kt_renderer::kRing, hipHostMallocDefault), "pinned params");
    KT_CREATE_HIP(hipMalloc(&r->d_params, sizeof(kt::RenderParams) * kt_renderer::kRing), "hipMalloc params");
    for (int i = 0; i < kt_renderer::kRing; ++i)
        KT_CREATE_HIP(hipEventCreateWithFlags(&r->param_ev[i], hipEventDisableTiming), "param event");
    // frame buffers (init.rs:142-178), allocated for both frames
    const size_t px = (size_t)cfg->width * cfg->height;
    for (auto& f : r->frames) {
        KT_CREATE_HIP(hipMalloc(&f.hdr, px * 4 * sizeof(float)), "hipMalloc HDR");
        KT_CREATE_HIP(hipMalloc(&f.bloom, px * 4 * sizeof(float)), "hipMalloc bloom");
        KT_CREATE_HIP(hipMalloc(&f.ldr, px * sizeof(uint32_t)), "hipMalloc image");
        KT_CREATE_HIP(hipMalloc(&f.flag, 2 * sizeof(uint32_t)), "hipMalloc LUT flag");
        KT_CREATE_HIP(hipHostMalloc(&f.ldr_host, px * sizeof(uint32_t), hipHostMallocDefault), "pinned image");
        KT_CREATE_HIP(hipHostMalloc(&f.flag_host, 2 * sizeof(uint32_t), hipHostMallocDefault), "pinned flag");
        f.flag_host[0] = f.flag_host[1] = 0;
        KT_CREATE_HIP(hipEventCreateWithFlags(&f.ready, hipEventDisableTiming), "frame event");
    }
#undef KT_CREATE_HIP
    *out = r;
    return KT_OK;
}

void kt_destroy(kt_renderer* r) {
    if (!r) return;
    for (auto& f : r->frames)
        if (f.stream) (void)hipStreamSynchronize(f.stream);
    for (auto& sl : r->slots)
        if (sl.used) (void)hipEventSynchronize(sl.done);
    release(r);
}

kt_status kt_get_derived(const kt_renderer* r, kt_derived* out) {
    if (!r || !out) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    fill_derived(r->d, out);
    return KT_OK;
}

kt_status kt_get_luts(const kt_renderer* r, double* disk, double* bb) {
    if (!r) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null renderer");
    if (disk) std::memcpy(disk, r->d.disk_lut.data(), r->d.disk_lut.size() * sizeof(double));
    if (bb) std::memcpy(bb, r->d.bb_lut.data(), r->d.bb_lut.size() * sizeof(double));
    return KT_OK;
}

kt_status kt_set_launch_opts(kt_renderer* r, const kt_launch_opts* o) {
    if (!r || !o) return set_err(r, KT_ERR_INVALID_ARGUMENT, "null argument");
    if (o->waves_per_cu < 0 || o->waves_per_cu > 32 || o->refill_threshold < 0 || o->refill_threshold > 64 ||
        o->queue_chunk < 0 || o->queue_chunk > 256 || o->cost_order < 0 || o->cost_order > 2 || o->long_ray_attempts < 0 ||
        o->isolate_level < 0 || o->isolate_level > 3)
        return set_err(r, KT_ERR_INVALID_ARGUMENT, "launch option out of range");
    r->opts = *o;
    return KT_OK;
}

kt_status kt_trace(kt_renderer* r, const kt_camera* cam, float* hdr, uint8_t* cls, void* stream) {
    if (!r || !cam || !hdr) return set_err(r, KT_ERR_INVALID_ARGUMENT, "null argument");
    return trace_into(r, cam, hdr, cls, nullptr, (hipStream_t)stream);
}

kt_status kt_trace_ex(kt_renderer* r, const kt_camera* cam, float* hdr, uint8_t* cls, uint32_t* attempts,
                      void* stream) {
    if (!r || !cam || !hdr) return set_err(r, KT_ERR_INVALID_ARGUMENT, "null argument");
    return trace_into(r, cam, hdr, cls, nullptr, (hipStream_t)stream, attempts);
}

kt_status kt_get_trace_stats(kt_renderer* r, kt_trace_stats* out) {
    if (!r || !out) return set_err(r, KT_ERR_INVALID_ARGUMENT, "null argument");
    if (r->last_slot < 0) return set_err(r, KT_ERR_INVALID_ARGUMENT, "no trace has been issued");
    const kt_renderer::Slot& sl = r->slots[r->last_slot];
    KT_HIP(r, hipEventSynchronize(sl.done), "trace sync");
    kt::DevCounters c;
    uint32_t flag[2];
    KT_HIP(r, hipMemcpy(&c, sl.counters, sizeof c, hipMemcpyDeviceToHost), "read counters");
    KT_HIP(r, hipMemcpy(flag, sl.flag_used, sizeof flag, hipMemcpyDeviceToHost), "read flag");
    out->rays = sl.rays;
    out->rhs_evals = c.rhs_evals;
    out->attempts = c.attempts;
    out->accepted = c.accepted;
    out->crossings = c.crossings;
    out->max_attempts_ray = c.max_attempts_ray;
    out->lut_overflow = flag[0];
    float T;
    std::memcpy(&T, &flag[1], sizeof T);
    out->overflow_temperature = flag[0] ? T : 0.0f;
    // device wall clock ticks -> microseconds
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, r->device) != hipSuccess || rate_khz <= 0)
        rate_khz = 100000;
    const double us_per_tick = 1000.0 / (double)rate_khz;
    out->kernel_us = c.clock_end > c.clock_start ? (double)(c.clock_end - c.clock_start) * us_per_tick : 0.0;
    out->tail_us = (c.clock_drain != ~0ull && c.clock_end > c.clock_drain)
                       ? (double)(c.clock_end - c.clock_drain) * us_per_tick
                       : 0.0;
    float ms = 0.0f;
    out->kernel_event_ms = hipEventElapsedTime(&ms, sl.ev0, sl.done) == hipSuccess ? (double)ms : 0.0;
    return KT_OK;
}

kt_status kt_post_process(kt_renderer* r, const float* hdr, float* bloom, uint32_t* ldr, void* stream) {
    if (!r || !hdr || !ldr || (r->d.bloom_active && !bloom)) return set_err(r, KT_ERR_INVALID_ARGUMENT, "null argument");
    return post_into(r, hdr, bloom, ldr, (hipStream_t)stream);
}

kt_status kt_post_process_noalloc(kt_renderer* r, const float* hdr, uint32_t* ldr, void* stream) {
    if (!r || !hdr || !ldr) return set_err(r, KT_ERR_INVALID_ARGUMENT, "null argument");
    if (!r->d.bloom_active) return post_into(r, hdr, nullptr, ldr, (hipStream_t)stream);
    const size_t px = (size_t)r->cfg.width * r->cfg.height;
    if (px > kt::noalloc_max_pixels())
        return set_err(r, KT_ERR_CAPACITY,
                       "kt_post_process_noalloc: a " + std::to_string(r->cfg.width) + "x" +
                           std::to_string(r->cfg.height) + " frame exceeds the library's bloom array of " +
                           std::to_string(kt::noalloc_max_pixels()) + " pixels (use kt_post_process)");
    if (!r->d_noalloc_bloom)
        return set_err(r, KT_ERR_DEVICE, "kt_post_process_noalloc: bloom array unavailable: " + r->noalloc_err);
    return post_into(r, hdr, r->d_noalloc_bloom, ldr, (hipStream_t)stream);
}

kt_status kt_submit_render_f64(kt_renderer* r, const kt_camera* cam, int32_t* submitted) {
    if (!r || !cam || !submitted) return set_err(r, KT_ERR_INVALID_ARGUMENT, "null argument");
    *submitted = 0;
    if (r->queue.size() == 2) return KT_OK;  // both frames in flight (mod.rs:127-129)
    int idx = -1;
    for (int off = 0; off < 2; ++off) {
        int i = (r->next_frame + off) % 2;
        if (!r->frames[i].pending) { idx = i; break; }
    }
    if (idx < 0) return set_err(r, KT_ERR_DEVICE, "no free frame buffer");
    auto& f = r->frames[idx];
    // Each frame gets its own stream, so two submitted frames overlap on the
    // device.  They are created on first use: a process has few hardware
    // queues (four by default) and streams beyond that share them.
    if (!f.stream) KT_HIP(r, hipStreamCreateWithFlags(&f.stream, hipStreamNonBlocking), "hipStreamCreate");
    hipStream_t s = f.stream;
    int rc = trace_into(r, cam, f.hdr, nullptr, f.flag, s);  // clears f.flag first
    if (rc != KT_OK) return rc;
    KT_HIP(r, hipMemcpyAsync(f.flag_host, f.flag, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "flag copy");
    rc = post_into(r, f.hdr, f.bloom, f.ldr, s);
    if (rc != KT_OK) return rc;
    size_t px = (size_t)r->cfg.width * r->cfg.height;
    KT_HIP(r, hipMemcpyAsync(f.ldr_host, f.ldr, px * sizeof(uint32_t), hipMemcpyDeviceToHost, s), "image copy");
    KT_HIP(r, hipEventRecord(f.ready, s), "frame event");
    f.pending = true;
    r->queue.push_back(idx);
    r->next_frame = (idx + 1) % 2;
    *submitted = 1;
    return KT_OK;
}

kt_status kt_submit_render(kt_renderer* r, const float cam_pos[3], const float fwd[3], const float rgt[3],
                           const float up[3], float fov_scale, int32_t* submitted) {
    if (!cam_pos || !fwd || !rgt || !up) return set_err(r, KT_ERR_INVALID_ARGUMENT, "null argument");
    kt_camera c;
    for (int k = 0; k < 3; ++k) {
        c.pos[k] = cam_pos[k];
        c.fwd[k] = fwd[k];
        c.rgt[k] = rgt[k];
        c.up[k] = up[k];
    }
    c.fov_scale = fov_scale;
    return kt_submit_render_f64(r, &c, submitted);
}

kt_status kt_ready_frame(kt_renderer* r, kt_frame_view* view, int32_t* has_frame) {
    if (!r || !view || !has_frame) return set_err(r, KT_ERR_INVALID_ARGUMENT, "null argument");
    *has_frame = 0;
    if (r->queue.empty()) return KT_OK;
    int idx = r->queue.front();
    auto& f = r->frames[idx];
    hipError_t q = hipEventQuery(f.ready);
    if (q == hipErrorNotReady) return KT_OK;
    if (q != hipSuccess) return hip_err(r, q, "frame event query");
    if (f.flag_host[0] != 0) {
        // the frame stays queued as in mod.rs:236-241; its index is reported
        // so the caller can still release it with kt_finish_frame
        view->index = idx;
        view->data = nullptr;
        float T;
        std::memcpy(&T, &f.flag_host[1], sizeof T);
        char b[160];
        std::snprintf(b, sizeof b, "colour temperature exceeds lut_max_temp: %g > %g", (double)T, r->cfg.lut_max_temp);
        return set_err(r, KT_ERR_LUT_OVERFLOW, b);
    }
    view->index = idx;
    view->width = r->cfg.width;
    view->height = r->cfg.height;
    view->reserved = 0;
    view->data = f.ldr_host;
    *has_frame = 1;
    return KT_OK;
}

kt_status kt_finish_frame(kt_renderer* r, int32_t index) {
    if (!r) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null renderer");
    if (r->queue.empty()) return set_err(r, KT_ERR_QUEUE_ORDER, "frame queue is empty");
    int expect = r->queue.front();
    if (expect != index) {
        char b[96];
        std::snprintf(b, sizeof b, "frame queue order: expect=%d got=%d", expect, index);
        return set_err(r, KT_ERR_QUEUE_ORDER, b);
    }
    r->queue.pop_front();
    r->frames[index].pending = false;
    return KT_OK;
}


kt_status kt_batch_create(const kt_config* cfgs, const kt_camera* cams, int32_t n, kt_batch** out) {
    if (!cfgs || !cams || !out || n <= 0) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument or n <= 0");
    *out = nullptr;
    if (n > 65536) return set_err(nullptr, KT_ERR_CAPACITY, "more than 65536 renders in one batch");
    kt_batch* b = new (std::nothrow) kt_batch();
    if (!b) return set_err(nullptr, KT_ERR_DEVICE, "out of host memory");
    b->n = n;
    b->cfgs.assign(cfgs, cfgs + n);
    b->cams.assign(cams, cams + n);
    b->ds.resize(n);
    b->precision = cfgs[0].precision;
    uint64_t px = 0;
    for (int i = 0; i < n; ++i) {
        std::string why;
        int rc = derive(b->cfgs[i], b->ds[i], why);
        if (rc != KT_OK) {
            release_batch(b);
            return set_err(nullptr, rc, "render " + std::to_string(i) + ": " + why);
        }
        if (cfgs[i].precision != b->precision) {
            release_batch(b);
            return set_err(nullptr, KT_ERR_INVALID_CONFIG, "all renders of a batch must share the precision mode");
        }
        px += (uint64_t)cfgs[i].width * cfgs[i].height;
    }
#define KT_BATCH_HIP(call, what)                  \
    do {                                          \
        hipError_t e_ = (call);                   \
        if (e_ != hipSuccess) {                   \
            int rc_ = hip_err(nullptr, e_, what); \
            release_batch(b);                     \
            return rc_;                           \
        }                                         \
    } while (0)
    KT_BATCH_HIP(hipGetDevice(&b->device), "hipGetDevice");
    hipDeviceProp_t prop;
    KT_BATCH_HIP(hipGetDeviceProperties(&prop, b->device), "hipGetDeviceProperties");
    b->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    b->d_disk_luts.assign(n, nullptr);
    b->d_bb_luts.assign(n, nullptr);
    for (int i = 0; i < n; ++i) {
        const size_t m = cfgs[i].lut_size;
        KT_BATCH_HIP(hipMalloc(&b->d_disk_luts[i], m * sizeof(double)), "hipMalloc disk LUT");
        KT_BATCH_HIP(hipMemcpy(b->d_disk_luts[i], b->ds[i].disk_lut.data(), m * sizeof(double), hipMemcpyHostToDevice),
                     "upload disk LUT");
        // renders with the same blackbody table share one device copy
        int same = -1;
        for (int j = 0; j < i && same < 0; ++j)
            if (same_blackbody(cfgs[i], cfgs[j])) same = j;
        if (same >= 0) continue;
        KT_BATCH_HIP(hipMalloc(&b->d_bb_luts[i], m * 3 * sizeof(double)), "hipMalloc blackbody LUT");
        KT_BATCH_HIP(hipMemcpy(b->d_bb_luts[i], b->ds[i].bb_lut.data(), m * 3 * sizeof(double), hipMemcpyHostToDevice),
                     "upload blackbody LUT");
    }
    KT_BATCH_HIP(hipHostMalloc(&b->h_params, sizeof(kt::RenderParams) * n, hipHostMallocDefault), "pinned params");
    KT_BATCH_HIP(hipMalloc(&b->d_params, sizeof(kt::RenderParams) * n), "hipMalloc params");
    KT_BATCH_HIP(hipMalloc(&b->d_counters, sizeof(kt::DevCounters)), "hipMalloc counters");
    KT_BATCH_HIP(hipMalloc(&b->d_flag, 2 * sizeof(uint32_t)), "hipMalloc flag");
    KT_BATCH_HIP(hipMemset(b->d_flag, 0, 2 * sizeof(uint32_t)), "clear flag");
    // tile order buffers for tile edges >= 2
    size_t tiles = 0;
    for (int i = 0; i < n; ++i) tiles += (size_t)((cfgs[i].width + 1) / 2) * ((cfgs[i].height + 1) / 2);
    b->max_tiles = tiles;
    KT_BATCH_HIP(hipMalloc(&b->d_keys, tiles), "hipMalloc cost keys");
    KT_BATCH_HIP(hipMalloc(&b->d_hist, kt::cost_order_hist_words((int)tiles) * sizeof(uint32_t)),
                 "hipMalloc cost histogram");
    KT_BATCH_HIP(hipMalloc(&b->d_perm, tiles * sizeof(uint32_t)), "hipMalloc tile order");
    KT_BATCH_HIP(hipEventCreate(&b->ev0), "event");
    KT_BATCH_HIP(hipEventCreate(&b->ev1), "event");
    KT_BATCH_HIP(hipEventCreateWithFlags(&b->ev_params, hipEventDisableTiming), "event");
#undef KT_BATCH_HIP
    (void)px;
    *out = b;
    return KT_OK;
}

void kt_batch_destroy(kt_batch* b) {
    if (!b) return;
    if (b->last_stream) (void)hipStreamSynchronize(b->last_stream);
    release_batch(b);
}

kt_status kt_batch_set_launch_opts(kt_batch* b, const kt_launch_opts* o) {
    if (!b || !o) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    if (o->waves_per_cu < 0 || o->waves_per_cu > 32 || o->refill_threshold < 0 || o->refill_threshold > 64 ||
        o->queue_chunk < 0 || o->queue_chunk > 256 || o->queue_chunk == 1 || o->cost_order < 0 || o->cost_order > 2 || o->long_ray_attempts < 0 ||
        o->isolate_level < 0 || o->isolate_level > 3)
        return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "launch option out of range (batch tile edge >= 2)");
    b->opts = *o;
    return KT_OK;
}

kt_status kt_batch_trace(kt_batch* b, float* const* hdr, uint8_t* const* cls, void* stream) {
    if (!b || !hdr) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t s = (hipStream_t)stream;
    const int tile = b->opts.queue_chunk > 0 ? b->opts.queue_chunk : 8;
    // the pinned records of the previous call must have reached the device,
    // and its launch (on any stream) must be done with the device records,
    // counters and tile order this call rewrites
    if (b->params_in_flight) {
        hipError_t e = hipEventSynchronize(b->ev_params);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, b->ev1, 0);
        if (e != hipSuccess) return hip_err(nullptr, e, "previous trace wait");
    }
    uint32_t tile_begin = 0;
    uint64_t rays = 0;
    for (int i = 0; i < b->n; ++i) {
        if (!hdr[i]) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null HDR pointer");
        kt::RenderParams& P = b->h_params[i];
        make_params(nullptr, b->cfgs[i], b->ds[i], b->cams[i], hdr[i], cls ? cls[i] : nullptr, b->d_flag, tile, P);
        P.long_ray = b->opts.long_ray_attempts > 0 ? (uint32_t)b->opts.long_ray_attempts : 256u;
        P.isolate = (uint32_t)b->opts.isolate_level;
        P.disk_lut = b->d_disk_luts[i];
        int bb = i;
        while (!b->d_bb_luts[bb]) --bb;  // the first render with the same table owns it
        P.bb_lut = b->d_bb_luts[bb];
        P.tile_begin = tile_begin;
        tile_begin += P.n_tiles;
        rays += (uint64_t)b->cfgs[i].width * b->cfgs[i].height * b->cfgs[i].spp;
    }
    const bool order = b->opts.cost_order != 2 && tile <= 8 && tile_begin <= b->max_tiles;
    for (int i = 0; i < b->n; ++i) b->h_params[i].tile_perm = order ? b->d_perm : nullptr;
    hipError_t e = hipMemcpyAsync(b->d_params, b->h_params, sizeof(kt::RenderParams) * b->n, hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_err(nullptr, e, "param upload");
    e = hipEventRecord(b->ev_params, s);
    if (e != hipSuccess) return hip_err(nullptr, e, "param event");
    b->params_in_flight = true;
    e = kt::launch_reset(b->d_counters, b->d_flag, s);
    if (e != hipSuccess) return set_err(nullptr, KT_ERR_LAUNCH, std::string("reset launch: ") + hipGetErrorString(e));
    if (order) {
        e = kt::launch_cost_order(b->d_params, b->n, (int)tile_begin, b->d_keys, b->d_hist, b->d_perm, s);
        if (e != hipSuccess) return set_err(nullptr, KT_ERR_LAUNCH, std::string("cost order: ") + hipGetErrorString(e));
    }
    const uint64_t items = (uint64_t)tile_begin * (uint64_t)tile * (uint64_t)tile;
    int wpc = b->opts.waves_per_cu > 0 ? b->opts.waves_per_cu : 8;
    int blocks = std::max(1, (b->cus * wpc + 3) / 4);
    int thr = b->opts.refill_threshold > 0 ? std::min(b->opts.refill_threshold, 64) : 8;
    e = hipEventRecord(b->ev0, s);
    if (e != hipSuccess) return hip_err(nullptr, e, "trace event");
    bool fp32_rhs = true;  // one kernel for the batch: FP64 if any render needs it
    for (int i = 0; i < b->n; ++i) fp32_rhs = fp32_rhs && uses_fp32_rhs(b->cfgs[i]);
    // a batch of one is a single render: its own right-hand-side form
    const bool exact1 = b->cfgs[0].rhs_mode == KT_RHS_HAMILTONIAN;
    e = fp32_rhs ? kt::launch_trace_mixed(b->d_params, b->n, items, thr, b->d_counters, blocks, exact1, s)
                 : kt::launch_trace_fp64(b->d_params, b->n, items, thr, b->d_counters, blocks, exact1, s);
    if (e != hipSuccess) return set_err(nullptr, KT_ERR_LAUNCH, std::string("batch trace launch: ") + hipGetErrorString(e));
    e = hipEventRecord(b->ev1, s);
    if (e != hipSuccess) return hip_err(nullptr, e, "trace event");
    b->last_stream = s;
    b->last_rays = rays;
    return KT_OK;
}

kt_status kt_batch_get_stats(kt_batch* b, kt_trace_stats* out) {
    if (!b || !out) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    hipError_t e = hipStreamSynchronize(b->last_stream);
    if (e != hipSuccess) return hip_err(nullptr, e, "stream sync");
    kt::DevCounters c;
    uint32_t flag[2];
    e = hipMemcpy(&c, b->d_counters, sizeof c, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(flag, b->d_flag, sizeof flag, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_err(nullptr, e, "read counters");
    std::memset(out, 0, sizeof *out);
    out->rays = b->last_rays;
    out->rhs_evals = c.rhs_evals;
    out->attempts = c.attempts;
    out->accepted = c.accepted;
    out->crossings = c.crossings;
    out->max_attempts_ray = c.max_attempts_ray;
    out->lut_overflow = flag[0];
    float T;
    std::memcpy(&T, &flag[1], sizeof T);
    out->overflow_temperature = flag[0] ? T : 0.0f;
    int rate_khz = 0;
    if (hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, b->device) != hipSuccess || rate_khz <= 0)
        rate_khz = 100000;
    const double us_per_tick = 1000.0 / (double)rate_khz;
    out->kernel_us = c.clock_end > c.clock_start ? (double)(c.clock_end - c.clock_start) * us_per_tick : 0.0;
    out->tail_us = (c.clock_drain != ~0ull && c.clock_end > c.clock_drain)
                       ? (double)(c.clock_end - c.clock_drain) * us_per_tick
                       : 0.0;
    float ms = 0.0f;
    out->kernel_event_ms = hipEventElapsedTime(&ms, b->ev0, b->ev1) == hipSuccess ? (double)ms : 0.0;
    return KT_OK;
}


kt_status kt_grad_create(const kt_config* cfg, const kt_grad_opts* opts, kt_grad** out) {
    if (!cfg || !opts || !out) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    if (cfg->spp != 1) return set_err(nullptr, KT_ERR_INVALID_CONFIG, "the differentiable renderer needs spp = 1");
    if (cfg->rhs_mode != KT_RHS_HAMILTONIAN)
        return set_err(nullptr, KT_ERR_INVALID_CONFIG,
                       "the differentiable renderer needs the exact (Hamiltonian) right-hand side");
    if (!(opts->fov_deg > 0.0 && opts->fov_deg < 180.0) || opts->checkpoint_interval < 1 ||
        opts->checkpoint_interval > 256 || opts->log_steps_per_ray < 1 ||
        (opts->adjoint_precision != KT_ADJOINT_FP64 && opts->adjoint_precision != KT_ADJOINT_FP32))
        return set_err(nullptr, KT_ERR_INVALID_ARGUMENT,
                       "fov in (0, 180), checkpoint interval 1..256, log size >= 1, adjoint precision FP64 or FP32");
    kt_grad* g = new (std::nothrow) kt_grad();
    if (!g) return set_err(nullptr, KT_ERR_DEVICE, "out of host memory");
    g->cfg = *cfg;
    g->opts = *opts;
    std::string why;
    int rc = derive(g->cfg, g->d, why);
    if (rc != KT_OK) {
        release_grad(g);
        return set_err(nullptr, rc, why);
    }
    const uint64_t n_pix = (uint64_t)cfg->width * cfg->height;
    g->K = (uint32_t)opts->checkpoint_interval;
    g->stride = 8 + 2 * g->K;
    const uint64_t per_ray = 1 + ((uint64_t)opts->log_steps_per_ray + g->K - 1) / g->K;
    const uint64_t cap = n_pix * per_ray;
    if (n_pix > 0x3FFFFFFFull || cap >= 0xFFFFFFF0ull) {
        release_grad(g);
        return set_err(nullptr, KT_ERR_CAPACITY, "frame or step log too large");
    }
    g->n_pix = (uint32_t)n_pix;
    g->capacity = (uint32_t)cap;
#define KT_GRAD_HIP(call, what)                   \
    do {                                          \
        hipError_t e_ = (call);                   \
        if (e_ != hipSuccess) {                   \
            int rc_ = hip_err(nullptr, e_, what); \
            release_grad(g);                      \
            return rc_;                           \
        }                                         \
    } while (0)
    KT_GRAD_HIP(hipGetDevice(&g->device), "hipGetDevice");
    hipDeviceProp_t prop;
    KT_GRAD_HIP(hipGetDeviceProperties(&prop, g->device), "hipGetDeviceProperties");
    g->cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    const size_t n = cfg->lut_size;
    KT_GRAD_HIP(hipMalloc(&g->d_disk_lut, n * sizeof(double)), "hipMalloc disk LUT");
    KT_GRAD_HIP(hipMalloc(&g->d_dlut, n * sizeof(double)), "hipMalloc disk LUT tangent");
    KT_GRAD_HIP(hipMalloc(&g->d_bb_lut, 3 * n * sizeof(double)), "hipMalloc blackbody LUT");
    KT_GRAD_HIP(hipMemcpy(g->d_bb_lut, g->d.bb_lut.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice),
                "upload blackbody LUT");
    KT_GRAD_HIP(hipHostMalloc(&g->h_params, sizeof(kt::RenderParams), hipHostMallocDefault), "pinned params");
    KT_GRAD_HIP(hipHostMalloc(&g->h_pool, 2 * sizeof(uint32_t), hipHostMallocDefault), "pinned pool flag");
    KT_GRAD_HIP(hipMalloc(&g->d_params, sizeof(kt::RenderParams)), "hipMalloc params");
    KT_GRAD_HIP(hipMalloc(&g->d_counters, sizeof(kt::DevCounters)), "hipMalloc counters");
    KT_GRAD_HIP(hipMalloc(&g->d_flag, 2 * sizeof(uint32_t)), "hipMalloc flag");
    KT_GRAD_HIP(hipMalloc(&g->d_cls, n_pix), "hipMalloc classes");
    const size_t tiles = (size_t)((cfg->width + 7) / 8) * ((cfg->height + 7) / 8);
    KT_GRAD_HIP(hipMalloc(&g->d_keys, tiles), "hipMalloc cost keys");
    KT_GRAD_HIP(hipMalloc(&g->d_hist, kt::cost_order_hist_words((int)tiles) * sizeof(uint32_t)),
                "hipMalloc cost histogram");
    KT_GRAD_HIP(hipMalloc(&g->d_perm, tiles * sizeof(uint32_t)), "hipMalloc tile order");
    KT_GRAD_HIP(hipMalloc(&g->d_arena, (size_t)cap * g->stride * sizeof(double)), "hipMalloc step log");
    KT_GRAD_HIP(hipMalloc(&g->d_head, n_pix * sizeof(uint32_t)), "hipMalloc log heads");
    KT_GRAD_HIP(hipMalloc(&g->d_nsteps, n_pix * sizeof(uint32_t)), "hipMalloc log steps");
    KT_GRAD_HIP(hipMalloc(&g->d_pk, 2 * n_pix * sizeof(double)), "hipMalloc pt/kt");
    KT_GRAD_HIP(hipMalloc(&g->d_pool, 2 * sizeof(uint32_t)), "hipMalloc pool head");
    KT_GRAD_HIP(hipMalloc(&g->d_grad_px, 4 * n_pix * sizeof(double)), "hipMalloc per-pixel gradient");
    g->threads = (uint32_t)std::max(256, (g->cus * 8 + 3) / 4 * 256);
    const size_t sc = (size_t)(6 * g->K + 6) * g->threads;
    KT_GRAD_HIP(hipMalloc(&g->d_scratch, sc * sizeof(double)), "hipMalloc adjoint scratch");
    KT_GRAD_HIP(hipMalloc(&g->d_queue, sizeof(unsigned long long)), "hipMalloc queue");
    KT_GRAD_HIP(hipMalloc(&g->d_rkeys, n_pix), "hipMalloc adjoint keys");
    KT_GRAD_HIP(hipMalloc(&g->d_rhist, kt::cost_order_hist_words((int)n_pix) * sizeof(uint32_t)),
                "hipMalloc adjoint sort counters");
    KT_GRAD_HIP(hipMalloc(&g->d_rorder, n_pix * sizeof(uint32_t)), "hipMalloc adjoint order");
    KT_GRAD_HIP(hipMalloc(&g->d_part, 4 * 256 * sizeof(double)), "hipMalloc partial sums");
    for (auto& e : g->ev) KT_GRAD_HIP(hipEventCreate(&e), "event");
    KT_GRAD_HIP(hipEventCreateWithFlags(&g->ev_params, hipEventDisableTiming), "event");
#undef KT_GRAD_HIP
    *out = g;
    return KT_OK;
}

void kt_grad_destroy(kt_grad* g) {
    if (!g) return;
    if (g->have_forward) {
        (void)hipStreamSynchronize(g->fwd_stream);
        (void)hipStreamSynchronize(g->last_stream);
    }
    release_grad(g);
}

kt_status kt_grad_forward(kt_grad* g, const double theta[4], float* hdr, uint8_t* hit_class, void* stream) {
    if (!g || !theta || !hdr) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    for (int k = 0; k < 4; ++k)
        if (!std::isfinite(theta[k])) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "theta is not finite");
    if (!(theta[2] > 0.0) || !(theta[3] > 0.0))
        return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "r_obs and mdot must be positive");
    hipStream_t s = (hipStream_t)stream;
    // the previous frame's work reads the buffers rewritten below: its
    // forward (trace, class copy) and its backward, which may have run on
    // another stream
    if (g->have_forward) {
        hipError_t e = hipStreamSynchronize(g->fwd_stream);
        if (e == hipSuccess) e = hipStreamSynchronize(g->last_stream);
        if (e != hipSuccess) return hip_err(nullptr, e, "previous frame wait");
    }
    std::string why;
    int rc = derive_grad(g, theta, why);
    if (rc != KT_OK) return set_err(nullptr, rc, why);
    const size_t n = g->cfg.lut_size;
    hipError_t e = hipMemcpy(g->d_disk_lut, g->d.disk_lut.data(), n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(g->d_dlut, g->dlut.data(), n * sizeof(double), hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_err(nullptr, e, "LUT upload");
    kt_config c = g->cfg;
    c.spin = theta[0];
    c.mdot_scale = theta[3];
    c.escape_radius = theta[2];
    kt_camera cam;
    const double si = std::sin(theta[1]), ci = std::cos(theta[1]);
    const double pos[3] = {theta[2] * si, theta[2] * ci, 0.0}, fwd[3] = {-si, -ci, 0.0};
    const double rgt[3] = {0.0, 0.0, -1.0}, up[3] = {-ci, si, 0.0};
    for (int k = 0; k < 3; ++k) {
        cam.pos[k] = pos[k];
        cam.fwd[k] = fwd[k];
        cam.rgt[k] = rgt[k];
        cam.up[k] = up[k];
    }
    cam.fov_scale = std::tan(g->opts.fov_deg * kPi / 360.0);
    kt::RenderParams& P = *g->h_params;
    const int tile = 8;
    make_params(nullptr, c, g->d, cam, hdr, g->d_cls, g->d_flag, tile, P);
    P.long_ray = 256u;
    P.disk_lut = g->d_disk_lut;
    P.bb_lut = g->d_bb_lut;
    P.tile_perm = g->d_perm;
    P.log.arena = g->d_arena;
    P.log.head = g->d_head;
    P.log.nsteps = g->d_nsteps;
    P.log.pk = g->d_pk;
    P.log.pool_head = g->d_pool;
    P.log.overflow = g->d_pool + 1;
    P.log.capacity = g->capacity;
    P.log.K = g->K;
    P.log.stride = g->stride;
    g->items = (uint64_t)P.n_tiles * tile * tile;
    e = hipMemcpyAsync(g->d_params, g->h_params, sizeof(kt::RenderParams), hipMemcpyHostToDevice, s);
    if (e != hipSuccess) return hip_err(nullptr, e, "param upload");
    e = kt::launch_cost_order(g->d_params, 1, (int)P.n_tiles, g->d_keys, g->d_hist, g->d_perm, s);
    if (e == hipSuccess) e = kt::launch_reset(g->d_counters, g->d_flag, s);
    if (e == hipSuccess) e = kt::launch_reset_log(g->d_pool, g->d_pool + 1, g->n_pix, s);
    if (e != hipSuccess) return set_err(nullptr, KT_ERR_LAUNCH, std::string("forward set-up: ") + hipGetErrorString(e));
    const int blocks = std::max(1, (g->cus * 8 + 3) / 4);
    e = hipEventRecord(g->ev[0], s);
    if (e == hipSuccess)
        e = uses_fp32_rhs(g->cfg) ? kt::launch_trace_mixed_record(g->d_params, g->items, 8, g->d_counters, blocks, s)
                                  : kt::launch_trace_fp64_record(g->d_params, g->items, 8, g->d_counters, blocks, s);
    if (e == hipSuccess) e = hipEventRecord(g->ev[1], s);
    if (e != hipSuccess) return set_err(nullptr, KT_ERR_LAUNCH, std::string("recording trace: ") + hipGetErrorString(e));
    if (hit_class) {
        e = hipMemcpyAsync(hit_class, g->d_cls, g->n_pix, hipMemcpyDeviceToDevice, s);
        if (e != hipSuccess) return hip_err(nullptr, e, "class copy");
    }
    g->last_stream = s;
    g->fwd_stream = s;
    g->have_forward = true;
    g->have_backward = false;
    return KT_OK;
}

kt_status kt_grad_backward(kt_grad* g, const float* dl_dhdr, double* grad, double* grad_px, void* stream) {
    if (!g || !dl_dhdr || !grad) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    if (!g->have_forward) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "kt_grad_backward before kt_grad_forward");
    hipStream_t s = (hipStream_t)stream;
    // the forward pass may have run on another stream: order after it
    hipError_t e = hipStreamWaitEvent(s, g->ev[1], 0);
    if (e != hipSuccess) return hip_err(nullptr, e, "forward wait");
    // a step log that ran out of room cannot be reversed
    e = hipMemcpyAsync(g->h_pool, g->d_pool, 2 * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_err(nullptr, e, "step log check");
    if (g->h_pool[1] != 0u)
        return set_err(nullptr, KT_ERR_CAPACITY, "the step log overflowed: raise kt_grad_opts.log_steps_per_ray");
    kt::GradConsts G;
    G.dlut = g->d_dlut;
    G.d_rin_da = g->d_rin_da;
    G.dTs_da = g->dTs_da;
    G.dTs_dm = g->dTs_dm;
    G.incl = g->theta[1];
    G.r_obs = g->theta[2];
    G.fov_scale = std::tan(g->opts.fov_deg * kPi / 360.0);
    G.w_scale = g->cfg.exposure_scale * (1.0 / (double)g->cfg.spp);
    e = hipEventRecord(g->ev[2], s);
    if (e == hipSuccess)
        e = kt::launch_grad_backward(g->d_params, G, dl_dhdr, grad_px ? grad_px : g->d_grad_px, g->d_scratch,
                                     g->threads, g->d_queue, g->d_nsteps, g->d_rkeys, g->d_rhist, g->d_rorder,
                                     (int)g->cfg.width, (int)g->cfg.height, g->n_pix, g->d_part, grad,
                                     uses_fp32_rhs(g->cfg), g->opts.adjoint_precision == KT_ADJOINT_FP32, s);
    if (e == hipSuccess) e = hipEventRecord(g->ev[3], s);
    if (e != hipSuccess) return set_err(nullptr, KT_ERR_LAUNCH, std::string("adjoint: ") + hipGetErrorString(e));
    g->last_stream = s;
    g->have_backward = true;
    return KT_OK;
}

kt_status kt_grad_get_stats(kt_grad* g, kt_grad_stats* out) {
    if (!g || !out) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    std::memset(out, 0, sizeof *out);
    if (!g->have_forward) return KT_OK;
    hipError_t e = hipStreamSynchronize(g->last_stream);
    if (e != hipSuccess) return hip_err(nullptr, e, "stream sync");
    kt::DevCounters c;
    uint32_t pool[2];
    e = hipMemcpy(&c, g->d_counters, sizeof c, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(pool, g->d_pool, sizeof pool, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_err(nullptr, e, "read counters");
    float ms = 0.0f;
    out->forward_ms = hipEventElapsedTime(&ms, g->ev[0], g->ev[1]) == hipSuccess ? (double)ms : 0.0;
    if (g->have_backward)
        out->backward_ms = hipEventElapsedTime(&ms, g->ev[2], g->ev[3]) == hipSuccess ? (double)ms : 0.0;
    out->rhs_evals = c.rhs_evals;
    out->attempts = c.attempts;
    out->chunks = pool[0];
    out->log_overflow = pool[1];
    out->checkpoint_interval = g->K;
    return KT_OK;
}

// ------------------------------------------------------------- autotuner
// Candidate launch options: resident waves per CU (one or two per SIMD,
// or both mixed), idle lanes that trigger a refill (how many rays a wave
// carries before it takes new work), the tile edge of the work order, and
// whether a wave holding a long ray stops taking new rays.
static const kt_launch_opts kTuneGrid[] = {
    {8, 16, 8, 1, 256, 0}, {8, 32, 8, 1, 256, 0}, {4, 16, 8, 1, 256, 0}, {4, 32, 8, 1, 256, 0},
    {8, 8, 4, 1, 256, 0},  {4, 16, 4, 1, 256, 0}, {6, 32, 8, 1, 256, 1}, {8, 32, 8, 1, 256, 1},
    {4, 32, 8, 1, 256, 1}, {6, 16, 8, 1, 256, 1}, {8, 16, 8, 1, 256, 2}, {6, 32, 8, 1, 256, 2},
};
static const int kTuneReps = 2;        // first pass: best of two runs per candidate
static const int kTuneFinalists = 3;   // second pass: the fastest three ...
static const int kTuneFinalReps = 4;   // ... timed four more times each, lowest mean wins

}  // extern "C"

// Two-pass choice over kTuneGrid with time_one(opts, &ms) timing one
// trace.  The best-of-two first pass ranks every candidate cheaply; the
// mean of the second pass decides among the leaders, so one lucky run
// cannot pick options that are slower on average.
template <typename TimeOne>
int tune_two_pass(TimeOne time_one, kt_launch_opts* best) {
    const int n = (int)(sizeof kTuneGrid / sizeof kTuneGrid[0]);
    std::vector<std::pair<float, int>> rank;
    for (int i = 0; i < n; ++i) {
        float t = 3.0e38f;
        for (int k = 0; k < kTuneReps; ++k) {
            float ms = 0.0f;
            int rc = time_one(kTuneGrid[i], &ms);
            if (rc != KT_OK) return rc;
            t = std::min(t, ms);
        }
        rank.emplace_back(t, i);
    }
    std::sort(rank.begin(), rank.end());
    double best_mean = 3.0e38;
    int best_i = rank[0].second;
    for (int f = 0; f < std::min(kTuneFinalists, n); ++f) {
        const int i = rank[f].second;
        double sum = 0.0;
        for (int k = 0; k < kTuneFinalReps; ++k) {
            float ms = 0.0f;
            int rc = time_one(kTuneGrid[i], &ms);
            if (rc != KT_OK) return rc;
            sum += ms;
        }
        if (sum / kTuneFinalReps < best_mean) {
            best_mean = sum / kTuneFinalReps;
            best_i = i;
        }
    }
    *best = kTuneGrid[best_i];
    return KT_OK;
}

extern "C" {

kt_status kt_autotune(kt_renderer* r, const kt_camera* cam, float* hdr, void* stream, kt_launch_opts* best) {
    if (!r || !cam || !hdr) return set_err(r, KT_ERR_INVALID_ARGUMENT, "null argument");
    hipStream_t s = (hipStream_t)stream;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    KT_HIP(r, hipEventCreate(&e0), "event");
    hipError_t ce = hipEventCreate(&e1);
    if (ce != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_err(r, ce, "event");
    }
    const kt_launch_opts before = r->opts;
    kt_launch_opts best_o = before;
    int rc = tune_two_pass(
        [&](const kt_launch_opts& o, float* ms) {
            r->opts = o;
            hipError_t e = hipEventRecord(e0, s);
            if (e != hipSuccess) return hip_err(r, e, "autotune timing");
            int trc = trace_into(r, cam, hdr, nullptr, nullptr, s);
            if (trc != KT_OK) return trc;
            e = hipEventRecord(e1, s);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
            return e == hipSuccess ? (int)KT_OK : hip_err(r, e, "autotune timing");
        },
        &best_o);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    r->opts = rc == KT_OK ? best_o : before;
    if (best && rc == KT_OK) *best = best_o;
    return rc;
}

kt_status kt_batch_autotune(kt_batch* b, float* const* hdr, void* stream, kt_launch_opts* best) {
    if (!b || !hdr) return set_err(nullptr, KT_ERR_INVALID_ARGUMENT, "null argument");
    const kt_launch_opts before = b->opts;
    kt_launch_opts best_o = before;
    int rc = tune_two_pass(
        [&](const kt_launch_opts& o, float* ms) {
            b->opts = o;
            int trc = kt_batch_trace(b, hdr, nullptr, stream);
            if (trc != KT_OK) return trc;
            hipError_t e = hipEventSynchronize(b->ev1);
            if (e == hipSuccess) e = hipEventElapsedTime(ms, b->ev0, b->ev1);
            return e == hipSuccess ? (int)KT_OK : hip_err(nullptr, e, "autotune timing");
        },
        &best_o);
    b->opts = rc == KT_OK ? best_o : before;
    if (best && rc == KT_OK) *best = best_o;
    return rc;
}
}  // extern "C"
