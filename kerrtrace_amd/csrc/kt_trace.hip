// Persistent FP64 geodesic tracer for gfx950.
//
// The reference (cuda/kernel.cu:88-142) runs one thread per pixel on a
// 16x16 grid, so every block waits for its slowest photon-ring ray.  Here a
// grid of resident waves pulls work from one queue instead:
//
//  * one lane = one ray; its whole state (position, momentum, FSAL k1,
//    step size, colour, transmittance) lives in VGPRs for its lifetime;
//  * the loop body is ONE DP5(4) attempt (six right-hand sides), so lanes
//    that retry a rejected step and lanes that move on stay converged;
//  * when a lane's ray terminates it goes idle; once `refill_threshold`
//    lanes of the wave are idle the wave claims that many new work items
//    with a single atomic on the queue head and hands them to the idle
//    lanes (ballot + mbcnt), so terminated rays are compacted out without
//    moving any state;
//  * work items are ordered in square pixel tiles so a wave's rays start
//    coherent; a batched launch concatenates the work items of many renders
//    and each lane carries its render's hot parameters.
//
// Per ray the arithmetic is that of cuda/kerr_trace.cuh:4-240 in binary64
// (see kt_math.hpp for the reproducibility contract).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"
#include "kt_math.hpp"
#include "kt_tracer_common.hpp"

namespace kt {

template <bool BATCH>
__global__ __launch_bounds__(256) void trace_fp64_persistent(const RenderParams* __restrict__ Ps, int n_renders,
                                                             uint64_t total_items, int refill_threshold,
                                                             DevCounters* __restrict__ counters) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    Lane L;
    L.active = false;
    L.rid = 0;
    L.overflow = false;
    L.overflow_T = 0.0;
    L.ray_att = 0;
    LaneCounters cnt;
    bool drained = false;
    Hot H = load_hot(prm<BATCH>(Ps, 0));

    for (;;) {
        // ---- refill idle lanes from the work queue (wave-uniform branch)
        uint64_t idle = __ballot(!L.active);
        if (idle != 0 && !drained) {
            int n_idle = __popcll(idle);
            if (n_idle >= refill_threshold || idle == __ballot(1)) {
                int leader = __ffsll((unsigned long long)idle) - 1;
                unsigned long long base = 0;
                if (lane == leader) base = atomicAdd(&counters->queue_head, (unsigned long long)n_idle);
                base = __shfl(base, leader);
                if (base + (unsigned long long)n_idle >= total_items) drained = true;
                if (!L.active) {
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
                        L.active = true;
                        start_sample(L, P, H);
                        if (!(0 < H.max_steps && L.Tr > H.cutoff)) end_sample(L, P, H);
                    }
                }
            }
        }
        if (drained && __ballot(L.active) == 0) break;
        if (!L.active) continue;

        // ---- one DP5(4) attempt (cuda/kerr_trace.cuh:83-122)
        const Geo g = H.g;
        const V6 s = L.s;
        if (!L.have_k1) {
            L.k1 = rhs(s, L.pt, g);
            L.have_k1 = true;
            cnt.rhs++;
        }
        const V6 k1 = L.k1;
        const double pt = L.pt;
        const double hs = L.h;
        // stage inputs s + h * sum_j a_ij k_j as fma chains
#define KT_ST2(c) fma(hs, KT_A21 * k1.c, s.c)
#define KT_ST3(c) fma(hs, fma(KT_A32, k2.c, KT_A31 * k1.c), s.c)
#define KT_ST4(c) fma(hs, fma(KT_A43, k3.c, fma(KT_A42, k2.c, KT_A41 * k1.c)), s.c)
#define KT_ST5(c) fma(hs, fma(KT_A54, k4.c, fma(KT_A53, k3.c, fma(KT_A52, k2.c, KT_A51 * k1.c))), s.c)
#define KT_ST6(c) \
    fma(hs, fma(KT_A65, k5.c, fma(KT_A64, k4.c, fma(KT_A63, k3.c, fma(KT_A62, k2.c, KT_A61 * k1.c)))), s.c)
#define KT_ST7(c) \
    fma(hs, fma(KT_B6, k6.c, fma(KT_B5, k5.c, fma(KT_B4, k4.c, fma(KT_B3, k3.c, KT_B1 * k1.c)))), s.c)
#define KT_STAGE(M) V6{M(x), M(y), M(z), M(px), M(py), M(pz)}
        V6 k2 = rhs(KT_STAGE(KT_ST2), pt, g);
        V6 k3 = rhs(KT_STAGE(KT_ST3), pt, g);
        V6 k4 = rhs(KT_STAGE(KT_ST4), pt, g);
        V6 k5 = rhs(KT_STAGE(KT_ST5), pt, g);
        V6 k6 = rhs(KT_STAGE(KT_ST6), pt, g);
        // the 7th stage input is the 5th-order solution (a7j = b_j)
        const V6 ns = KT_STAGE(KT_ST7);
        V6 k7 = rhs(ns, pt, g);
        cnt.rhs += 6;
        // embedded error, max norm relative to the start state (kerr_trace.cuh:104-116)
        const double tol = L.tol;
#define KT_ERR(c)                                                                                               \
    (fabs(hs * fma(KT_E7, k7.c, fma(KT_E6, k6.c, fma(KT_E5, k5.c, fma(KT_E4, k4.c, fma(KT_E3, k3.c, KT_E1 * k1.c)))))) / \
     (tol * fmax(fabs(s.c), 1.0)))
        double ex = KT_ERR(x), ey = KT_ERR(y), ez = KT_ERR(z);
        double epx = KT_ERR(px), epy = KT_ERR(py), epz = KT_ERR(pz);
#undef KT_ST2
#undef KT_ST3
#undef KT_ST4
#undef KT_ST5
#undef KT_ST6
#undef KT_ST7
#undef KT_STAGE
#undef KT_ERR
        double err = fmax(fmax(fmax(ex, ey), fmax(ez, epx)), fmax(epy, epz));
        cnt.attempts++;
        L.ray_att++;
        L.att_in_step++;
        const bool accepted = err <= 1.0;
        // controller without clamps (kerr_trace.cuh:119)
        L.h = hs * 0.9 * pow_m02(err);
        if (!accepted && L.att_in_step < H.max_attempts) continue;  // retry from the same state

        // ---- commit; a rejected last attempt is committed too (kerr_trace.cuh:123-134)
        if (accepted) cnt.accepted++;
        L.s = ns;
        L.att_in_step = 0;
        double r_now = ks_r(ns.x, ns.y, ns.z, g.aa);
        bool done = false;
        const RenderParams& P = prm<BATCH>(Ps, L.rid);
        if (r_now < H.rh_eps) {  // horizon, tested before the disk (kerr_trace.cuh:135-137)
            L.cls = (L.cls & ~3) | 1;
            done = true;
        } else {
            if ((s.y < 0.0 && ns.y >= 0.0) || (s.y > 0.0 && ns.y <= 0.0)) {
                // equatorial crossing on the dense output (kerr_trace.cuh:138-174)
                cnt.crossings++;
                double t;
                if (P.disk_root_mode == 1) {
                    t = disk_root(s.y, ns.y, k1.y, k3.y, k4.y, k5.y, k6.y, k7.y, hs);
                } else {
                    double den = s.y - ns.y;
                    t = den != 0.0 ? s.y / den : 0.5;
                    t = fmin(fmax(t, 0.0), 1.0);
                }
                double x_hit = dense(s.x, ns.x, k1.x, k3.x, k4.x, k5.x, k6.x, k7.x, hs, t);
                double z_hit = dense(s.z, ns.z, k1.z, k3.z, k4.z, k5.z, k6.z, k7.z, hs, t);
                double px_hit = dense(s.px, ns.px, k1.px, k3.px, k4.px, k5.px, k6.px, k7.px, hs, t);
                double py_hit = dense(s.py, ns.py, k1.py, k3.py, k4.py, k5.py, k6.py, k7.py, hs, t);
                double pz_hit = dense(s.pz, ns.pz, k1.pz, k3.pz, k4.pz, k5.pz, k6.pz, k7.pz, hs, t);
                if (shade_crossing(P, x_hit, z_hit, px_hit, py_hit, pz_hit, L.pt, L.kt, L.col, L.Tr, L.overflow,
                                   L.overflow_T)) {
                    L.cls |= 4;
                    if (((L.cls >> 5) & 7) < 7) L.cls += 32;  // contributions, saturating at 7
                }
            }
            if (r_now > H.R_esc) {  // escape: sky along normalize(dx/dl) (kerr_trace.cuh:229-237)
                bool grid = false;
                double sk = sky(P, ns, L.pt, L.Tr, grid);
                cnt.rhs++;
                L.col[0] = L.col[0] + sk;
                L.col[1] = L.col[1] + sk;
                L.col[2] = L.col[2] + sk;
                L.cls = (L.cls & ~3) | 2 | (grid ? 8 : 0);
                done = true;
            } else if (!(__builtin_isfinite(ns.x) && __builtin_isfinite(ns.y) && __builtin_isfinite(ns.z) &&
                         __builtin_isfinite(ns.px) && __builtin_isfinite(ns.py) && __builtin_isfinite(ns.pz))) {
                done = true;  // a non-finite state can add no further colour
            }
        }
        if (accepted) L.k1 = k7;  // FSAL
        L.have_k1 = accepted;
        if (!done) {
            L.step += 1;
            if (L.step < H.max_steps && L.Tr > H.cutoff) {
                L.tol = step_tolerance(L.s, r_now, g, H.tol, H.tol_floor);
            } else {
                done = true;
            }
        }
        if (done) {
            cnt.max_att = max(cnt.max_att, L.ray_att);
            end_sample(L, P, H);
        }
    }

    // ---- wave-level reduction of the work counters, one atomic per wave
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
    }
}

__global__ void reset_counters(DevCounters* c, uint32_t* error_flag) {
    if (threadIdx.x == 0) {
        c->rhs_evals = 0;
        c->attempts = 0;
        c->accepted = 0;
        c->crossings = 0;
        c->max_attempts_ray = 0;
        c->queue_head = 0;
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

// d_params holds n_renders records sorted by ray_begin (one for a single render).
hipError_t launch_trace_fp64(const RenderParams* d_params, int n_renders, uint64_t total_items, int refill_threshold,
                             DevCounters* d_counters, int blocks, hipStream_t stream) {
    if (n_renders == 1)
        hipLaunchKernelGGL(trace_fp64_persistent<false>, dim3(blocks), dim3(256), 0, stream, d_params, 1,
                           total_items, refill_threshold, d_counters);
    else
        hipLaunchKernelGGL(trace_fp64_persistent<true>, dim3(blocks), dim3(256), 0, stream, d_params, n_renders,
                           total_items, refill_threshold, d_counters);
    return hipGetLastError();
}

}  // namespace kt
