// Device pieces shared by the FP64 and the mixed-precision persistent
// tracers: DP5(4) tableau, potential-scaled tolerance, lookup-table fetches,
// disk shading, sky, the work-queue decode and the per-lane ray lifecycle.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"
#include "kt_math.hpp"

namespace kt {


// DP5(4) tableau (cuda/kerr_trace.cuh:52-59, docs/dormand_prince_extended_tableau.md)
#define KT_A21 (1.0 / 5.0)
#define KT_A31 (3.0 / 40.0)
#define KT_A32 (9.0 / 40.0)
#define KT_A41 (44.0 / 45.0)
#define KT_A42 (-56.0 / 15.0)
#define KT_A43 (32.0 / 9.0)
#define KT_A51 (19372.0 / 6561.0)
#define KT_A52 (-25360.0 / 2187.0)
#define KT_A53 (64448.0 / 6561.0)
#define KT_A54 (-212.0 / 729.0)
#define KT_A61 (9017.0 / 3168.0)
#define KT_A62 (-355.0 / 33.0)
#define KT_A63 (46732.0 / 5247.0)
#define KT_A64 (49.0 / 176.0)
#define KT_A65 (-5103.0 / 18656.0)
#define KT_B1 (35.0 / 384.0)
#define KT_B3 (500.0 / 1113.0)
#define KT_B4 (125.0 / 192.0)
#define KT_B5 (-2187.0 / 6784.0)
#define KT_B6 (11.0 / 84.0)
#define KT_E1 (35.0 / 384.0 - 5179.0 / 57600.0)
#define KT_E3 (500.0 / 1113.0 - 7571.0 / 16695.0)
#define KT_E4 (125.0 / 192.0 - 393.0 / 640.0)
#define KT_E5 (-2187.0 / 6784.0 + 92097.0 / 339200.0)
#define KT_E6 (11.0 / 84.0 - 187.0 / 2100.0)
#define KT_E7 (-1.0 / 40.0)

static constexpr int kWave = 64;

struct Lane;
// A wave holding long-running rays (photon-ring orbits) issues ahead of
// the other wave(s) on its SIMD, so the serial chain of steps of the
// frame's longest rays, which bounds the frame time, is slowed as little
// as possible: priority 1 past `long_ray` attempts, 2 past 3x, 3 past 8x.
__device__ __forceinline__ void set_wave_priority(int p) {
    switch (p) {
        case 0: __builtin_amdgcn_s_setprio(0); break;
        case 1: __builtin_amdgcn_s_setprio(1); break;
        case 2: __builtin_amdgcn_s_setprio(2); break;
        default: __builtin_amdgcn_s_setprio(3); break;
    }
}

// Potential-scaled tolerance of the step that starts at s, whose
// Kerr-Schild radius is r_c (cuda/kerr_trace.cuh:67-73, with the tolerance
// floor extension).  The radius is the one the previous commit computed
// for its horizon and escape tests, so it is not solved again.
__device__ __forceinline__ double step_tolerance(const V6& s, double r_c, const Geo& g, double tol,
                                                 double tol_floor) {
    double r2_c = r_c * r_c;
    double dH = fma(r2_c, r2_c, g.aa * s.y * s.y);
    double Hc = g.M * r_c * r2_c * div_pos(1.0, dH);
    double tq = (Hc - 0.05) / (0.3 - 0.05);
    tq = fmin(fmax(tq, 0.0), 1.0);
    return tol * fmax(5.0 - 5.0 * tq, tol_floor);
}

// Blackbody colour of temperature T (cuda/kernel.cu:1-16); the texture's
// clamped linear filter is done with exact weights at texel coordinate
// T / T_max * (N - 1).
__device__ __forceinline__ void blackbody(const RenderParams& P, double T, double rgb[3], bool& overflow) {
    if (T > P.lut_max_temp) {
        overflow = true;
        rgb[0] = rgb[1] = rgb[2] = 0.0;
        return;
    }
    double t_norm = T / P.lut_max_temp;
    double pos = t_norm * (double)(P.lut_n - 1);
    double fl = floor(pos);
    int i0 = (int)fl;
    double w = pos - fl;
    int i1 = i0 + 1;
    i0 = min(max(i0, 0), P.lut_n - 1);
    i1 = min(max(i1, 0), P.lut_n - 1);
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double c0 = P.bb_lut[i0 * 3 + c], c1 = P.bb_lut[i1 * 3 + c];
        rgb[c] = c0 + w * (c1 - c0);
    }
}

// Disk flux LUT at u in [0,1] (cuda/kerr_trace.cuh:196-198): normalised
// texture coordinate without a half-texel offset, i.e. texel u*N - 0.5.
__device__ __forceinline__ double disk_factor(const RenderParams& P, double u) {
    double xb = u * (double)P.lut_n - 0.5;
    double fl = floor(xb);
    int i0 = (int)fl;
    double w = xb - fl;
    int i1 = i0 + 1;
    i0 = min(max(i0, 0), P.lut_n - 1);
    i1 = min(max(i1, 0), P.lut_n - 1);
    double c0 = P.disk_lut[i0], c1 = P.disk_lut[i1];
    return c0 + w * (c1 - c0);
}

// Emission of one equatorial crossing (cuda/kerr_trace.cuh:174-226).  The
// square roots and quotients whose operands are positive by construction
// (a hit on the disk lies outside the ISCO, so r > |a|, r^2 + a^2 > 0, the
// disk span and LUT step are positive) use the guard-free cores; the
// others (signed angular momentum, the redshift factor) keep the library
// division.  The
// lane's colour and transmittance stay in the caller's registers: the
// function returns the colour to add, d_col * alpha * Tr, and the factor
// 1 - alpha the caller multiplies into Tr (passing lane state by pointer
// into a non-inlined call would move the whole lane into scratch memory).
struct Shade {
    double r, g, b;       // colour to add
    double tr_mul;        // transmittance factor
    double overflow_T;    // temperature above lut_max_temp, or 0
    int composited;
};
static __device__ __noinline__ Shade shade_crossing(const RenderParams& P, double x_hit, double z_hit, double px_hit,
                                                    double py_hit, double pz_hit, double pt, double kt, double Tr) {
    Shade o{0.0, 0.0, 0.0, 1.0, 0.0, 0};
    const double a = P.a, M = P.M, aa = P.aa;
    double r_hit = ks_r(x_hit, 0.0, z_hit, aa);
    if (!(r_hit >= P.disk_inner && r_hit <= P.disk_outer)) return o;
    double sqrt_M = sqrt_pos(M), r_sqrt = sqrt_pos(r_hit);
    double disc_denom = 1.0 - 3.0 * M / r_hit + 2.0 * a * sqrt_M / (r_hit * r_sqrt);
    if (!(disc_denom > 0.0)) return o;
    double omega = -sqrt_M / (r_hit * r_sqrt + a * sqrt_M);
    double r2_hit = r_hit * r_hit;
    double denom_hit = r2_hit + aa;
    double inv_dh = div_pos(1.0, denom_hit);
    double hlx = (r_hit * x_hit - a * z_hit) * inv_dh;
    double hly = 0.0;
    double hlz = (r_hit * z_hit + a * x_hit) * inv_dh;
    double ldotp = fma(hlx, px_hit, fma(hly, py_hit, hlz * pz_hit));
    double ldotxi = hlx * z_hit - hlz * x_hit;
    double dHh = r2_hit * r2_hit;
    double Hh = M * r_hit * r2_hit * div_pos(1.0, dHh);
    double L = (z_hit * px_hit - x_hit * pz_hit) + 2.0 * Hh * ldotxi * (ldotp + kt);
    double energy = -pt;
    double l_over_e = L / energy;
    double g = 1.0 / (div_pos(1.0, sqrt_pos(disc_denom)) * (1.0 - omega * l_over_e));
    double dspan = P.disk_outer - P.disk_inner;
    double uu = dspan > 0.0 ? div_pos(r_hit - P.disk_inner, dspan) : 0.0;
    double sampled = disk_factor(P, uu);
    double T = P.T_scale * sampled * g;
    if (P.noise_detail > 0) {
        // Boyer-Lindquist phi of the hit (cuda/kerr_math.cuh:60-81; r_hit is
        // overwritten with r2 * inv_r as in kerr_trace.cuh:203)
        double rho2 = fma(x_hit, x_hit, fma(0.0, 0.0, z_hit * z_hit));
        double u = rho2 - aa;
        double s_sq = fma(u, u, 4.0 * aa * 0.0 * 0.0);
        double inv_s = div_pos(1.0, sqrt_pos(s_sq));
        double s = s_sq * inv_s;
        double r2 = 0.5 * (u + s);
        double inv_r = div_pos(1.0, sqrt_pos(r2));
        r_hit = r2 * inv_r;
        double cos_th = 0.0 * inv_r;
        double cos2 = cos_th * cos_th;
        double sin_sq = 1.0 - cos2;
        double inv_sin = div_pos(1.0, sqrt_pos(sin_sq));
        double denom = r2 + aa;
        double inv = inv_sin * div_pos(1.0, denom);
        double cos_phi = (r_hit * z_hit + a * x_hit) * inv;
        double sin_phi = (r_hit * x_hit - a * z_hit) * inv;
        double phi = atan2(sin_phi, cos_phi);
        double r3 = r_hit * r_hit * r_hit;
        double om = div_pos(1.0, a + sqrt_pos(r3 * P.inv_M));
        double phase = phi + P.noise_winding * om;
        double sn, cs;
        sincos(phase, &sn, &cs);
        double nu = r_hit * cs * P.noise_scale;
        double nv = r_hit * sn * P.noise_scale;
        double nval = fbm(nu, nv, P.noise_detail, P.noise_seed_mix);
        double modulator = 1.0 + P.noise_strength * (nval - 0.5);
        T = T * modulator;
    }
    if (!(__builtin_isfinite(T) && T > 0.0)) return o;
    double d_col[3];
    bool ovf = false;
    blackbody(P, T, d_col, ovf);
    if (ovf) o.overflow_T = T;
    double luma = 0.2126 * d_col[0] + 0.7152 * d_col[1] + 0.0722 * d_col[2];
    double luma_scaled = div_pos(luma, P.wl_step);
    double alpha = 1.0 - exp(-luma_scaled * 0.5);
    o.r = d_col[0] * alpha * Tr;
    o.g = d_col[1] * alpha * Tr;
    o.b = d_col[2] * alpha * Tr;
    o.tr_mul = 1.0 - alpha;
    o.composited = alpha > 0.0 ? 1 : 0;
    return o;
}

// Applies a crossing's emission to the lane's colour, transmittance and
// class (same operations as cuda/kerr_trace.cuh:221-224).
__device__ __forceinline__ void apply_shade(const Shade& sh, double col[3], double& Tr, int& cls, bool& overflow,
                                            double& overflow_T) {
    col[0] = col[0] + sh.r;
    col[1] = col[1] + sh.g;
    col[2] = col[2] + sh.b;
    Tr = Tr * sh.tr_mul;
    if (sh.overflow_T > 0.0 && !overflow) {
        overflow = true;
        overflow_T = sh.overflow_T;
    }
    if (sh.composited) {
        cls |= 4;
        if (((cls >> 5) & 7) < 7) cls += 32;  // contributions, saturating at 7
    }
}

// Sky grid intensity along the escape direction (cuda/kerr_math.cuh:157-171);
// the state arrives by value for the same reason as in shade_crossing.
struct Sky {
    double intensity;
    int grid;
};
static __device__ __noinline__ Sky sky(const RenderParams& P, double x, double y, double z, double px, double py,
                                       double pz, double pt, double Tr) {
    const Geo g = make_geo(P.a, P.aa, P.M, P.rhs_exact != 0);
    V6 ke = rhs(V6{x, y, z, px, py, pz}, pt, g);
    double q = ke.x * ke.x + ke.y * ke.y + ke.z * ke.z;
    double inv = 1.0 / sqrt(q);
    double dx = ke.x * inv, dy = ke.y * inv, dz = ke.z * inv;
    double theta = atan2(dz, dx);
    double phi_lat = asin(dy);
    double spacing = P.sky_spacing;
    double lt = P.sky_lt;
    double d_theta = fmod(fabs(theta), spacing);
    double d_phi = fmod(fabs(phi_lat), spacing);
    bool grid = (d_theta < lt || d_theta > (spacing - lt)) || (d_phi < lt || d_phi > (spacing - lt)) ||
                (fabs(phi_lat) < lt * 3.0);
    return Sky{grid ? P.sky_int * Tr : 0.0, grid ? 1 : 0};
}

struct LaneCounters {
    uint32_t rhs = 0, attempts = 0, accepted = 0, crossings = 0, max_att = 0;
};

// Decode a global work item into (render, pixel).  Items past the end of a
// render's tiled range or outside its image return false.
// The parameter record of render rid (always record 0 for a single render,
// so its loads are wave-uniform scalar loads).
template <bool BATCH>
__device__ __forceinline__ const RenderParams& prm(const RenderParams* __restrict__ Ps, int rid) {
    return BATCH ? Ps[rid] : Ps[0];
}

// Render owning global tile g (records sorted by tile_begin).
template <bool BATCH>
__device__ __forceinline__ int render_of_tile(const RenderParams* __restrict__ Ps, int n_renders, uint32_t g) {
    if (!BATCH) return 0;
    int lo = 0, hi = n_renders - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (Ps[mid].tile_begin <= g) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Decode a work item: item = slot * T^2 + within, where slot is the visiting
// position of a global tile (tile_perm maps it to the tile; row-major
// without it) and within the pixel inside the tile.  Pixels of edge tiles
// that fall outside their image return false.
template <bool BATCH>
__device__ __forceinline__ bool decode_item(const RenderParams* __restrict__ Ps, int n_renders, uint64_t item,
                                            int& rid, int& px, int& py) {
    const uint32_t T = (uint32_t)Ps[0].tile;
    const uint32_t tt = T * T;
    const uint32_t slot = (uint32_t)(item / tt);
    const uint32_t within = (uint32_t)(item - (uint64_t)slot * tt);
    const uint32_t g = Ps[0].tile_perm ? Ps[0].tile_perm[slot] : slot;
    rid = render_of_tile<BATCH>(Ps, n_renders, g);
    const RenderParams& P = prm<BATCH>(Ps, rid);
    const uint32_t tile_id = g - P.tile_begin;
    if (tile_id >= P.n_tiles) return false;
    const uint32_t ty = tile_id / (uint32_t)P.tiles_x;
    const uint32_t tx = tile_id - ty * (uint32_t)P.tiles_x;
    px = (int)(tx * T + within % T);
    py = (int)(ty * T + within / T);
    return px < P.W && py < P.H;
}

// Per-lane ray state of the persistent tracer.
struct Lane {
    V6 s, k1;
    double pt, kt, h, tol, Tr;
    double col[3], acc[3];
    double overflow_T;
    int rid, pix_x, pix_y, sample, step, att_in_step, cls;
    double hx, hz, hpx, hpy, hpz;  // dense-output state at the equatorial crossing
    double r_now;     // Kerr-Schild radius of the committed state
    uint32_t ray_att;
    uint32_t chunk, nrec;  // recording traces: current log chunk, committed steps
    bool active, finished, have_k1, overflow;
    bool post;        // the committed step still needs its crossing/escape/budget pass
    bool crossing;    // ... and it crossed the equator
};

// Hot per-render values a lane keeps in registers.
struct Hot {
    Geo g;
    double tol, tol_floor, cutoff, rh_eps, R_esc;
    int max_steps, max_attempts, spp;
};

// A wave-uniform double moved from an SGPR pair into a VGPR pair (an empty
// asm whose operand must sit in VGPRs).  The tracer's scalar registers are
// the scarce ones (stage coefficients, exec masks of nested branches); the
// thresholds that a committed step reads once cost nothing in VGPRs, while
// spilled SGPRs cost a v_readlane per use on every step.
__device__ __forceinline__ double in_vgpr(double x) {
    asm("" : "+v"(x));
    return x;
}
__device__ __forceinline__ void hot_to_vgprs(Hot& H) {
    H.tol = in_vgpr(H.tol);
    H.tol_floor = in_vgpr(H.tol_floor);
    H.cutoff = in_vgpr(H.cutoff);
    H.rh_eps = in_vgpr(H.rh_eps);
    H.R_esc = in_vgpr(H.R_esc);
}

__device__ __forceinline__ Hot load_hot(const RenderParams& P) {
    Hot o;
    o.g = make_geo(P.a, P.aa, P.M, P.rhs_exact != 0);
    o.tol = P.tol;
    o.tol_floor = P.tol_floor;
    o.cutoff = P.cutoff;
    o.rh_eps = P.rh_eps;
    o.R_esc = P.R_esc;
    o.max_steps = P.max_steps;
    o.max_attempts = P.max_attempts;
    o.spp = P.spp;
    return o;
}

// Initial state of sample L.sample of the lane's pixel: jittered pinhole
// direction (cuda/kernel.cu:112-121) and the null momentum in Kerr-Schild
// form (cuda/kerr_trace.cuh:16-47).
__device__ __forceinline__ void start_sample(Lane& L, const RenderParams& P, const Hot& H) {
    const Geo& g = H.g;
    uint32_t pixel_index = (uint32_t)(L.pix_y * P.W + L.pix_x);
    uint64_t base = ((uint64_t)pixel_index << 32) | (uint64_t)(uint32_t)L.sample;
    double rx = (double)pcg32(base) * 0x1p-32;
    double ry = (double)pcg32(base ^ 0xda3e39cb94b95bdbULL) * 0x1p-32;
    double u = (((double)L.pix_x + rx) * P.inv_w_2 - 1.0) * P.aspect * P.fov_scale;
    double v = (1.0 - ((double)L.pix_y + ry) * P.inv_h_2) * P.fov_scale;
    double dx = u * P.rgt[0] + v * P.up[0] + P.fwd[0];
    double dy = u * P.rgt[1] + v * P.up[1] + P.fwd[1];
    double dz = u * P.rgt[2] + v * P.up[2] + P.fwd[2];
    double dd = dx * dx + dy * dy + dz * dz;
    double inv_n = 1.0 / sqrt(dd);
    double nx = dx * inv_n, ny = dy * inv_n, nz = dz * inv_n;
    double x = P.cam[0], y = P.cam[1], z = P.cam[2];
    double a2 = g.aa;
    double rho2 = fma(x, x, fma(y, y, z * z));
    double uu = rho2 - a2;
    double s_term = sqrt(fma(uu, uu, 4.0 * a2 * y * y));
    double r2 = 0.5 * (uu + s_term);
    double r = sqrt(r2);
    double denom = r2 + a2;
    double inv_denom = 1.0 / denom;
    double lx = (r * x - g.a * z) * inv_denom;
    double ly = y / r;
    double lz = (r * z + g.a * x) * inv_denom;
    double denomH = fma(r2, r2, a2 * y * y);
    double Hs = g.M * r2 * r * (1.0 / denomH);
    double ldotn = fma(lx, nx, fma(ly, ny, lz * nz));
    double g_tt = -1.0 + 2.0 * Hs;
    double g_tn = 2.0 * Hs * ldotn;
    double g_nn = 1.0 + 2.0 * Hs * ldotn * ldotn;
    double disc = g_tn * g_tn - g_tt * g_nn;
    L.kt = (-g_tn - sqrt(disc)) / g_tt;
    L.pt = g_tt * L.kt + g_tn;
    double sc = L.kt + ldotn;
    L.s = V6{x, y, z, nx + 2.0 * Hs * lx * sc, ny + 2.0 * Hs * ly * sc, nz + 2.0 * Hs * lz * sc};
    L.col[0] = L.col[1] = L.col[2] = 0.0;
    L.Tr = 1.0;
    L.h = P.h0;
    L.have_k1 = false;
    L.step = 0;
    L.att_in_step = 0;
    L.cls = 0;
    L.tol = step_tolerance(L.s, ks_r(L.s.x, L.s.y, L.s.z, g.aa), g, H.tol, H.tol_floor);
}

// A sample ended: accumulate it, then start the pixel's next sample (a
// sample whose step budget is empty is marked finished again and ends on
// the next pass) or write the pixel (cuda/kernel.cu:132-141) and free the
// lane.
__device__ __forceinline__ void end_sample(Lane& L, const RenderParams& P, const Hot& H) {
    L.acc[0] = L.acc[0] + L.col[0];
    L.acc[1] = L.acc[1] + L.col[1];
    L.acc[2] = L.acc[2] + L.col[2];
    L.sample += 1;
    if (L.sample < H.spp) {
        start_sample(L, P, H);
        L.finished = !(0 < H.max_steps && L.Tr > H.cutoff);
        return;
    }
    double fs = P.exposure_over_spp;
    size_t idx = (size_t)L.pix_y * (size_t)P.W + (size_t)L.pix_x;
    float4 out = make_float4((float)(L.acc[0] * fs), (float)(L.acc[1] * fs), (float)(L.acc[2] * fs), 1.0f);
    *reinterpret_cast<float4*>(P.hdr + idx * 4) = out;
    int c = L.cls | (L.overflow ? 16 : 0);
    if (P.cls) P.cls[idx] = (uint8_t)c;
    if (P.attempts_out) P.attempts_out[idx] = L.ray_att;
    if (L.overflow && P.error_flag) {
        if (atomicCAS(P.error_flag, 0u, 1u) == 0u) P.error_flag[1] = __float_as_uint((float)L.overflow_T);
    }
    L.active = false;
}

__device__ __forceinline__ void wave_priority(const Lane& L, uint32_t long_ray, int& prio) {
    const uint32_t att = L.active ? L.ray_att : 0u;
    int want = 0;
    if (__ballot(att > long_ray) != 0) {
        want = 1;
        if (__ballot(att > 3u * long_ray) != 0) {
            want = 2;
            if (__ballot(att > 8u * long_ray) != 0) want = 3;
        }
    }
    if (want != prio) {
        set_wave_priority(want);
        prio = want;
    }
}

// Recording traces: log committed step L.nrec (start state s, step size
// hs, transmittance before the step); a new chunk (with s as its
// checkpoint) starts every K steps.  Pool chunks are claimed with one
// atomic per wave.
__device__ __forceinline__ void record_step(Lane& L, const RenderParams& P, const V6& s, double hs, int lane) {
    const GradLog& G = P.log;
    const uint32_t u = L.nrec % G.K;
    const bool fresh = u == 0;
    const bool need_pool = fresh && L.nrec > 0 && L.chunk != kNoChunk;
    const uint64_t want = __ballot(need_pool);
    uint32_t base = 0;
    if (want) {
        const int leader = __ffsll((unsigned long long)want) - 1;
        if (lane == leader) base = atomicAdd(G.pool_head, (uint32_t)__popcll(want));
        base = __shfl(base, leader);
    }
    if (fresh) {
        uint32_t c = kNoChunk;
        if (L.nrec == 0) {
            c = (uint32_t)(L.pix_y * P.W + L.pix_x);
            double* pk = G.pk + 2 * (size_t)c;
            pk[0] = L.pt;
            pk[1] = L.kt;
        } else if (need_pool) {
            c = base + (uint32_t)__popcll(want & ((1ull << lane) - 1ull));
            if (c >= G.capacity) {
                atomicOr(G.overflow, 1u);
                c = kNoChunk;
            }
        }
        if (c != kNoChunk) {
            double* ch = G.arena + (size_t)c * G.stride;
            ch[0] = s.x;
            ch[1] = s.y;
            ch[2] = s.z;
            ch[3] = s.px;
            ch[4] = s.py;
            ch[5] = s.pz;
            ch[6] = __longlong_as_double((long long)(L.nrec == 0 ? 0xFFFFFFFFull : (unsigned long long)L.chunk));
        }
        L.chunk = c;
    }
    if (L.chunk != kNoChunk) {
        double* ch = G.arena + (size_t)L.chunk * G.stride;
        ch[8 + u] = hs;
        ch[8 + G.K + u] = L.Tr;
    }
    L.nrec += 1;
}

// Recording traces: a ray ended; its log head and step count.
__device__ __forceinline__ void record_end(const Lane& L, const RenderParams& P) {
    const size_t idx = (size_t)L.pix_y * (size_t)P.W + (size_t)L.pix_x;
    P.log.head[idx] = L.chunk;
    P.log.nsteps[idx] = L.chunk == kNoChunk ? 0u : L.nrec;
}

// Second half of a commit (cuda/kerr_trace.cuh:138-237 after the horizon
// test): the equatorial crossing's emission, the escape test with the sky
// colour, the non-finite stop, then the step counter and the loop condition
// (kerr_trace.cuh:62) with the next step's tolerance.  It runs at the top of
// the tracer's next iteration, where no stage derivative is live.
__device__ __forceinline__ void post_commit(Lane& L, const RenderParams& P, const Hot& H, LaneCounters& cnt) {
    if (L.crossing) {
        cnt.crossings++;
        const Shade sh = shade_crossing(P, L.hx, L.hz, L.hpx, L.hpy, L.hpz, L.pt, L.kt, L.Tr);
        apply_shade(sh, L.col, L.Tr, L.cls, L.overflow, L.overflow_T);
    }
    const V6& ns = L.s;
    bool done = false;
    if (L.r_now > H.R_esc) {  // escape: sky along normalize(dx/dl) (kerr_trace.cuh:229-237)
        const Sky sk = sky(P, ns.x, ns.y, ns.z, ns.px, ns.py, ns.pz, L.pt, L.Tr);
        cnt.rhs++;
        L.col[0] = L.col[0] + sk.intensity;
        L.col[1] = L.col[1] + sk.intensity;
        L.col[2] = L.col[2] + sk.intensity;
        L.cls = (L.cls & ~3) | 2 | (sk.grid ? 8 : 0);
        done = true;
    } else if (!(__builtin_isfinite(ns.x) && __builtin_isfinite(ns.y) && __builtin_isfinite(ns.z) &&
                 __builtin_isfinite(ns.px) && __builtin_isfinite(ns.py) && __builtin_isfinite(ns.pz))) {
        done = true;  // a non-finite state can add no further colour
    }
    if (!done) {
        L.step += 1;
        if (L.step < H.max_steps && L.Tr > H.cutoff) {
            L.tol = step_tolerance(L.s, L.r_now, H.g, H.tol, H.tol_floor);
        } else {
            done = true;
        }
    }
    L.finished = done;
}

// First half of a commit (kerr_trace.cuh:123-174): the attempt's end point
// becomes the state (also after max_attempts rejections), then the horizon
// test, then -- on a change of sign of y -- the crossing parameter on the
// step's DOPRI5 dense output (the clamped secant as written, or the FP64
// Newton/bisection root) and the interpolated hit state.  k1, k7 and d
// (sum_i d_i k_i) are the step's dense-output ingredients; r_ns is the
// Kerr-Schild radius of ns.  The emission
// and the rest of the step are left to post_commit.
__device__ __forceinline__ void commit(Lane& L, const V6& s, const V6& ns, double r_ns, double hs, const V6& k1,
                                       const V6& k7, const V6& d, const Hot& H, int disk_root_mode) {
    L.s = ns;
    L.att_in_step = 0;
    L.r_now = r_ns;  // ks_r(ns), from the right-hand side at ns
    if (L.r_now < H.rh_eps) {  // horizon, tested before the disk (kerr_trace.cuh:135-137)
        L.cls = (L.cls & ~3) | 1;
        L.finished = true;
        return;
    }
    // (bitwise & and |: no short-circuit branches)
    L.crossing = ((s.y < 0.0) & (ns.y >= 0.0)) | ((s.y > 0.0) & (ns.y <= 0.0));
    if (L.crossing) {
        double t;
        if (disk_root_mode == 1) {
            t = disk_root_s(s.y, ns.y, k1.y, k7.y, d.y, hs);
        } else {
            double den = s.y - ns.y;
            t = den != 0.0 ? s.y / den : 0.5;
            t = fmin(fmax(t, 0.0), 1.0);
        }
        L.hx = dense_s(s.x, ns.x, k1.x, k7.x, d.x, hs, t);
        L.hz = dense_s(s.z, ns.z, k1.z, k7.z, d.z, hs, t);
        L.hpx = dense_s(s.px, ns.px, k1.px, k7.px, d.px, hs, t);
        L.hpy = dense_s(s.py, ns.py, k1.py, k7.py, d.py, hs, t);
        L.hpz = dense_s(s.pz, ns.pz, k1.pz, k7.pz, d.pz, hs, t);
    }
    L.post = true;
}

// Wave-level reduction of the work counters at kernel exit: one atomic per
// counter per wave, plus the wave's exit time.
__device__ __forceinline__ void flush_counters(const LaneCounters& cnt, DevCounters* __restrict__ counters,
                                               int lane) {
    unsigned long long r = cnt.rhs, a = cnt.attempts, c = cnt.accepted, x = cnt.crossings, m = cnt.max_att;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        r += __shfl_xor(r, off);
        a += __shfl_xor(a, off);
        c += __shfl_xor(c, off);
        x += __shfl_xor(x, off);
        unsigned long long mo = __shfl_xor(m, off);
        m = mo > m ? mo : m;
    }
    if (lane == 0) {
        atomicAdd(&counters->rhs_evals, r);
        atomicAdd(&counters->attempts, a);
        atomicAdd(&counters->accepted, c);
        atomicAdd(&counters->crossings, x);
        atomicMax(&counters->max_attempts_ray, m);
        atomicMax(&counters->clock_end, (unsigned long long)wall_clock64());
    }
}

// A flag every lane of the wave holds equally, read from the first active
// lane: the compiler then treats branches on it as wave-uniform (scalar
// branches, no exec masks), which it cannot prove for a flag carried round
// a loop through blocks where lanes diverge.
__device__ __forceinline__ bool uniform_flag(bool b) { return __builtin_amdgcn_readfirstlane((int)b) != 0; }

// Refill the wave's idle lanes from the work queue: one atomic per wave
// claims as many items as there are idle lanes once `refill_threshold` of
// them are idle (or all are), and ballot + popcount hand them out.  Items
// past the end or outside their image leave the lane idle.  A new ray whose
// step budget is empty is marked finished right away.  Sets `drained` once
// the claim reaches the end of the queue; the caller stops calling it then.
template <bool BATCH>
__device__ __forceinline__ void refill(Lane& L, Hot& H, const RenderParams* __restrict__ Ps, int n_renders,
                                       uint64_t total_items, int refill_threshold, DevCounters* __restrict__ counters,
                                       int lane, bool& drained) {
    uint64_t idle = __ballot(!L.active);
    if (idle == 0) return;
    int n_idle = __popcll(idle);
    if (n_idle < refill_threshold && idle != __ballot(1)) return;
    int leader = __ffsll((unsigned long long)idle) - 1;
    unsigned long long base = 0;
    if (lane == leader) base = atomicAdd(&counters->queue_head, (unsigned long long)n_idle);
    // the leader's claim to every lane through scalar registers, so the
    // drain test below is a scalar branch
    base = ((unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(base >> 32), leader) << 32) |
           (unsigned long long)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)base, leader);
    if (base + (unsigned long long)n_idle >= total_items) {
        drained = true;
        if (lane == leader) atomicMin(&counters->clock_drain, (unsigned long long)wall_clock64());
    }
    if (L.active) return;
    uint64_t item = base + (uint64_t)__popcll(idle & ((1ull << lane) - 1ull));
    int rid = 0, px = 0, py = 0;
    if (item < total_items && decode_item<BATCH>(Ps, n_renders, item, rid, px, py)) {
        L.rid = rid;
        L.pix_x = px;
        L.pix_y = py;
        const RenderParams& P = prm<BATCH>(Ps, rid);
        if (BATCH) H = load_hot(P);
        L.sample = 0;
        L.acc[0] = L.acc[1] = L.acc[2] = 0.0;
        L.overflow = false;
        L.ray_att = 0;
        L.chunk = kNoChunk;
        L.nrec = 0;
        L.active = true;
        L.post = false;
        start_sample(L, P, H);
        L.finished = !(0 < H.max_steps && L.Tr > H.cutoff);
    }
}

}  // namespace kt
