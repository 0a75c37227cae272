// The persistent tracer's loop, shared by the FP64 kernel (kt_trace.hip)
// and the mixed-precision kernel (kt_mixed.hip).
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
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"
#include "kt_math.hpp"
#include "kt_tracer_common.hpp"

namespace kt {

// The loop of a persistent tracer kernel (one lane = one ray, one loop
// turn = one DP5(4) attempt of every active lane).
// RECORD: the differentiable renderer's forward pass, which also writes
// the step log (RenderParams::log) the adjoint pass replays.
// MIXED: the mixed-precision tracer: the right-hand sides run in rhs32
// (FP64 radius solve, FP32 bulk); the stage algebra, state, error norm,
// controller, dense output and shading stay FP64.
// RHS: how the right-hand side's dr/dy weight (Geo::w2c, w2f) is chosen:
// from each lane's render (batched launches), or fixed for the launch, so
// the exact form's unit factor and the as-written form's 1/2 fold away.
enum : int { kRhsPerRender = 0, kRhsExact = 1, kRhsAsWritten = 2 };

// The DP5(4) coefficients from stage 5 on (stage, solution, error and
// dense-output weights).  The FP64 tracer reads them from constant memory
// on every loop turn (scalar loads, eight coefficients per load) instead of
// forming them from 64-bit immediates: the kernel has no scalar registers
// to keep them across loop turns, and forming them took two s_mov_b32 each
// per attempt, issue slots on the serial chain of the wave that holds the
// frame's longest ray.  Stages 2 to 4 keep their few immediates, so the
// loads issued at the top of an attempt arrive while those stages run
// instead of stalling it.  The mixed tracer keeps immediates throughout
// (measured: the loads made its 3840x2160 frames 1.3% slower).
struct Dp5Coef {
    double a51, a52, a53, a54, a61, a62, a63, a64, a65;
    double b1, b3, b4, b5, b6, e1, e3, e4, e5, e6, e7, d1, d3, d4, d5, d6, d7;
};
static __constant__ Dp5Coef kDp5 = {KT_A51, KT_A52, KT_A53, KT_A54, KT_A61, KT_A62, KT_A63, KT_A64, KT_A65,
                                    KT_B1, KT_B3, KT_B4, KT_B5, KT_B6, KT_E1, KT_E3, KT_E4, KT_E5,
                                    KT_E6, KT_E7, Dense::d1, Dense::d3, Dense::d4, Dense::d5,
                                    Dense::d6, Dense::d7};
static constexpr Dp5Coef kDp5Imm = {KT_A51, KT_A52, KT_A53, KT_A54, KT_A61, KT_A62, KT_A63, KT_A64, KT_A65,
                                    KT_B1, KT_B3, KT_B4, KT_B5, KT_B6, KT_E1, KT_E3, KT_E4, KT_E5,
                                    KT_E6, KT_E7, Dense::d1, Dense::d3, Dense::d4, Dense::d5,
                                    Dense::d6, Dense::d7};
typedef const __attribute__((address_space(4))) Dp5Coef* Dp5Ptr;

template <bool BATCH, bool RECORD, bool MIXED, int RHS>
__device__ __forceinline__ void trace_loop(const RenderParams* __restrict__ Ps, int n_renders, uint64_t total_items,
                                           int refill_threshold, DevCounters* __restrict__ counters) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    Lane L;
    L.active = false;
    L.finished = false;
    L.post = false;
    L.crossing = false;
    L.rid = 0;
    L.overflow = false;
    L.overflow_T = 0.0;
    L.ray_att = 0;
    L.chunk = kNoChunk;
    L.nrec = 0;
    LaneCounters cnt;
    bool drained = false;
    int prio = 0;
    uint32_t turn = 0;  // loop turns (wave-uniform)
    const uint32_t long_ray = prm<BATCH>(Ps, 0).long_ray;
    const int isolate = (int)prm<BATCH>(Ps, 0).isolate;
    Hot H = load_hot(prm<BATCH>(Ps, 0));
    if (!BATCH) hot_to_vgprs(H);  // (a batch's lanes load their own)
    if (lane == 0) atomicMin(&counters->clock_start, (unsigned long long)wall_clock64());

    for (;;) {
        // ---- second half of last iteration's commit, then rays that ended:
        // write the pixel or start its next sample.  Both run here, where no
        // stage derivative is live.
        if (L.post) {
            L.post = false;
            post_commit(L, prm<BATCH>(Ps, L.rid), H, cnt);
        }
        if (L.finished) {
            L.finished = false;
            cnt.max_att = max(cnt.max_att, L.ray_att);
            if (RECORD) record_end(L, prm<BATCH>(Ps, L.rid));
            end_sample(L, prm<BATCH>(Ps, L.rid), H);
        }
        // ---- refill idle lanes from the work queue (wave-uniform branch)
        // (a wave isolated around a long ray takes no new rays until it ends)
        if (!uniform_flag(drained) && (isolate == 0 || prio < isolate))
            refill<BATCH>(L, H, Ps, n_renders, total_items, refill_threshold, counters, lane, drained);

        if (uniform_flag(drained) && __ballot(L.active) == 0) break;
        // a wave holding a long-running ray (photon-ring orbits) issues
        // ahead of the other wave(s) on its SIMD, so its serial chain of
        // steps, which bounds the frame time, is not slowed by them; the
        // thresholds are hundreds of attempts apart, so every 8th turn is
        // often enough (the check costs three ballots)
        if ((++turn & 7u) == 0u) wave_priority(L, long_ray, prio);
        // (no divergent `continue` or `break` in this loop: lanes without
        // an attempt this turn skip the block, so the loop's only exit is
        // the wave-uniform one above and no exec masks are carried round)
        if (L.active && !L.finished) {
            // ---- one DP5(4) attempt (cuda/kerr_trace.cuh:83-122)
            Geo g = H.g;
            if (RHS != kRhsPerRender) g = make_geo(g.a, g.aa, g.M, RHS == kRhsExact);
            const V6 s = L.s;
            const G32 g32 = make_g32(g);
            const float ptf = (float)L.pt;
            const double pt = L.pt;
#define KT_RHS(st) (MIXED ? rhs32((st), ptf, g, g32) : rhs((st), pt, g))
            if (!L.have_k1) {
                L.k1 = KT_RHS(s);
                L.have_k1 = true;
                cnt.rhs++;
            }
            const double hs = L.h;
            // the table's address is opaque to the compiler on each turn, so
            // its loads are issued inside the attempt, not hoisted out of
            // the loop into registers it does not have
            Dp5Ptr C = (Dp5Ptr)&kDp5;
            if (!MIXED) asm volatile("" : "+s"(C));
#define KT_C(f) (MIXED ? kDp5Imm.f : C->f)
            V6 ns, eacc, dacc;
            const V6 k1s = L.k1;
            {
                const V6 k1 = L.k1;
                // stage inputs s + h * sum_j a_ij k_j as fma chains
#define KT_ST2(c) fma(hs, KT_A21 * k1.c, s.c)
#define KT_ST3(c) fma(hs, fma(KT_A32, k2.c, KT_A31 * k1.c), s.c)
#define KT_ST4(c) fma(hs, fma(KT_A43, k3.c, fma(KT_A42, k2.c, KT_A41 * k1.c)), s.c)
#define KT_ST5(c) fma(hs, fma(KT_C(a54), k4.c, fma(KT_C(a53), k3.c, fma(KT_C(a52), k2.c, KT_C(a51) * k1.c))), s.c)
#define KT_ST6(c) \
        fma(hs, fma(KT_C(a65), k5.c, fma(KT_C(a64), k4.c, fma(KT_C(a63), k3.c, fma(KT_C(a62), k2.c, KT_C(a61) * k1.c)))), s.c)
#define KT_STAGE(M) V6{M(x), M(y), M(z), M(px), M(py), M(pz)}
                const V6 k2 = KT_RHS(KT_STAGE(KT_ST2));
                const V6 k3 = KT_RHS(KT_STAGE(KT_ST3));
                const V6 k4 = KT_RHS(KT_STAGE(KT_ST4));
                const V6 k5 = KT_RHS(KT_STAGE(KT_ST5));
                const V6 in6 = KT_STAGE(KT_ST6);
                // The 5th-order solution b.k and the error (b - bhat).k are fma
                // chains in stage order; their leading terms are folded now so
                // k2..k5 are dead while stage 6 and 7 are evaluated.
#define KT_B4ACC(c) fma(KT_C(b5), k5.c, fma(KT_C(b4), k4.c, fma(KT_C(b3), k3.c, KT_C(b1) * k1.c)))
#define KT_E5ACC(c) fma(KT_C(e5), k5.c, fma(KT_C(e4), k4.c, fma(KT_C(e3), k3.c, KT_C(e1) * k1.c)))
                V6 bacc = KT_STAGE(KT_B4ACC);
                eacc = KT_STAGE(KT_E5ACC);
                // dense-output sum d1 k1 + d3 k3 + ... + d7 k7 as an fma chain (dense_sum)
#define KT_D5ACC(c) fma(KT_C(d5), k5.c, fma(KT_C(d4), k4.c, fma(KT_C(d3), k3.c, KT_C(d1) * k1.c)))
                dacc = KT_STAGE(KT_D5ACC);
                const V6 k6 = KT_RHS(in6);
#define KT_NS(c) fma(hs, fma(KT_C(b6), k6.c, bacc.c), s.c)
#define KT_EACC6(c) fma(KT_C(e6), k6.c, eacc.c)
                ns = KT_STAGE(KT_NS);  // = the 7th stage input (a7j = b_j)
                eacc = KT_STAGE(KT_EACC6);
#define KT_D6ACC(c) fma(KT_C(d6), k6.c, dacc.c)
                dacc = KT_STAGE(KT_D6ACC);
#undef KT_ST2
#undef KT_ST3
#undef KT_ST4
#undef KT_ST5
#undef KT_ST6
#undef KT_B4ACC
#undef KT_E5ACC
#undef KT_D5ACC
#undef KT_D6ACC
#undef KT_NS
#undef KT_EACC6
            }
            // k7 = f(ns); the FP64 right-hand side also yields the Kerr-Schild
            // radius of ns, which the commit's horizon and escape tests need
            double r_ns = 0.0;
            const V6 k7 = MIXED ? rhs32(ns, ptf, g, g32) : rhs(ns, pt, g, &r_ns);
            if (MIXED) r_ns = ks_r(ns.x, ns.y, ns.z, g.aa);
#undef KT_RHS
            cnt.rhs += 6;
#define KT_D7ACC(c) fma(KT_C(d7), k7.c, dacc.c)
            dacc = KT_STAGE(KT_D7ACC);
#undef KT_D7ACC
            // embedded error, max norm relative to the start state (kerr_trace.cuh:104-116)
            const double tol = L.tol;
#define KT_ERR(c) div_pos(fabs(hs * fma(KT_C(e7), k7.c, eacc.c)), tol * fmax(fabs(s.c), 1.0))
            double ex = KT_ERR(x), ey = KT_ERR(y), ez = KT_ERR(z);
            double epx = KT_ERR(px), epy = KT_ERR(py), epz = KT_ERR(pz);
#undef KT_ERR
#undef KT_STAGE
            double err = fmax(fmax(fmax(ex, ey), fmax(ez, epx)), fmax(epy, epz));
            if (__builtin_expect(!(err <= 1.79769313486231570e308), 0)) {
                // a non-finite step error: the library quotients (special values)
#define KT_ERRX(c) (fabs(hs * fma(KT_C(e7), k7.c, eacc.c)) / (tol * fmax(fabs(s.c), 1.0)))
                err = fmax(fmax(fmax(KT_ERRX(x), KT_ERRX(y)), fmax(KT_ERRX(z), KT_ERRX(px))),
                           fmax(KT_ERRX(py), KT_ERRX(pz)));
#undef KT_ERRX
#undef KT_C
            }
            cnt.attempts++;
            L.ray_att++;
            L.att_in_step++;
            const bool accepted = err <= 1.0;
            // controller without clamps (kerr_trace.cuh:119)
            L.h = hs * 0.9 * pow_m02(err);
            // a rejected attempt is retried from the same state, unless it was
            // the step's last allowed one
            if (accepted || L.att_in_step >= H.max_attempts) {
                // ---- commit; a rejected last attempt is committed too (kerr_trace.cuh:123-134)
                if (accepted) cnt.accepted++;
                if (accepted) L.k1 = k7;  // FSAL
                L.have_k1 = accepted;
                if (RECORD) record_step(L, prm<BATCH>(Ps, L.rid), s, hs, lane);
                commit(L, s, ns, r_ns, hs, k1s, k7, dacc, H, prm<BATCH>(Ps, L.rid).disk_root_mode);
            }
        }
    }

    flush_counters(cnt, counters, lane);
}

}  // namespace kt
