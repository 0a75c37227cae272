// Persistent mixed-precision geodesic tracer for gfx950.
//
// Same work queue, lane lifecycle, termination and shading as the FP64
// tracer (kt_trace.hip); the DP5(4) attempt differs.  With
// dx/dl = p - corr and dp/dl = kp (rhs_mixed), the stage inputs of the
// same Dormand-Prince method are rewritten so that the large, smooth part
// of the position update never passes through FP32:
//
//   p_i = p + h sum_j a_ij kp_j
//   x_i = x + h c_i p + h^2 sum_k (A^2)_ik kp_k - h sum_j a_ij corr_j
//
// (A^2)_ik = sum_j a_ij a_jk, c_i = sum_j a_ij.  The FP64 terms are
// x + h c_i p; the FP32 sums only hold O(M/r) gravitational pieces.  The
// embedded error estimate likewise reduces to FP32 sums of those pieces
// (sum_j (b_j - bhat_j) = 0 cancels p exactly):
//
//   e_x = h (h sum_k (dc A)_k kp_k - sum_j dc_j corr_j),  e_p = h sum_j dc_j kp_j
//
// so FP32 rounding of the right-hand side does not set a noise floor on
// the error norm even at tolerance 1e-9.  FP64 is kept for: the state and
// its accumulation, the Kerr-Schild radius solve, the error norm and step
// controller, the dense-output disk root (FP64 interpolant rebuilt from the
// step's FP32 pieces) and shading.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"
#include "kt_math.hpp"
#include "kt_tracer_common.hpp"

namespace kt {

// (A^2)_ik, (bA)_k and (dc A)_k of the DP5(4) tableau, folded in FP64 then
// rounded once to FP32.
#define KT_C2 (1.0 / 5.0)
#define KT_C3 (3.0 / 10.0)
#define KT_C4 (4.0 / 5.0)
#define KT_C5 (8.0 / 9.0)
#define F(x) ((float)(x))
#define Q31 F(KT_A32 * KT_A21)
#define Q41 F(KT_A42 * KT_A21 + KT_A43 * KT_A31)
#define Q42 F(KT_A43 * KT_A32)
#define Q51 F(KT_A52 * KT_A21 + KT_A53 * KT_A31 + KT_A54 * KT_A41)
#define Q52 F(KT_A53 * KT_A32 + KT_A54 * KT_A42)
#define Q53 F(KT_A54 * KT_A43)
#define Q61 F(KT_A62 * KT_A21 + KT_A63 * KT_A31 + KT_A64 * KT_A41 + KT_A65 * KT_A51)
#define Q62 F(KT_A63 * KT_A32 + KT_A64 * KT_A42 + KT_A65 * KT_A52)
#define Q63 F(KT_A64 * KT_A43 + KT_A65 * KT_A53)
#define Q64 F(KT_A65 * KT_A54)
#define BA1 F(KT_B3 * KT_A31 + KT_B4 * KT_A41 + KT_B5 * KT_A51 + KT_B6 * KT_A61)
#define BA2 F(KT_B3 * KT_A32 + KT_B4 * KT_A42 + KT_B5 * KT_A52 + KT_B6 * KT_A62)
#define BA3 F(KT_B4 * KT_A43 + KT_B5 * KT_A53 + KT_B6 * KT_A63)
#define BA4 F(KT_B5 * KT_A54 + KT_B6 * KT_A64)
#define BA5 F(KT_B6 * KT_A65)
#define EA1 F(KT_E3 * KT_A31 + KT_E4 * KT_A41 + KT_E5 * KT_A51 + KT_E6 * KT_A61 + KT_E7 * KT_B1)
#define EA2 F(KT_E3 * KT_A32 + KT_E4 * KT_A42 + KT_E5 * KT_A52 + KT_E6 * KT_A62)
#define EA3 F(KT_E4 * KT_A43 + KT_E5 * KT_A53 + KT_E6 * KT_A63 + KT_E7 * KT_B3)
#define EA4 F(KT_E5 * KT_A54 + KT_E6 * KT_A64 + KT_E7 * KT_B4)
#define EA5 F(KT_E6 * KT_A65 + KT_E7 * KT_B5)
#define EA6 F(KT_E7 * KT_B6)
// dense-output weights: (d A)_k = sum_j d_j a_jk (a_7k = b_k); sum_j d_j = 0
#define DD1 Dense::d1
#define DD3 Dense::d3
#define DD4 Dense::d4
#define DD5 Dense::d5
#define DD6 Dense::d6
#define DD7 Dense::d7
#define DA1 F(DD3 * KT_A31 + DD4 * KT_A41 + DD5 * KT_A51 + DD6 * KT_A61 + DD7 * KT_B1)
#define DA2 F(DD3 * KT_A32 + DD4 * KT_A42 + DD5 * KT_A52 + DD6 * KT_A62)
#define DA3 F(DD4 * KT_A43 + DD5 * KT_A53 + DD6 * KT_A63 + DD7 * KT_B3)
#define DA4 F(DD5 * KT_A54 + DD6 * KT_A64 + DD7 * KT_B4)
#define DA5 F(DD6 * KT_A65 + DD7 * KT_B5)
#define DA6 F(DD7 * KT_B6)

struct X64 {
    double x, y, z;
};
struct P32 {
    float x, y, z;
};

template <bool BATCH>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void trace_mixed_persistent(
    const RenderParams* __restrict__ Ps, int n_renders, uint64_t total_items, int refill_threshold,
    DevCounters* __restrict__ counters) {
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
    K32 k1m{0, 0, 0, 0, 0, 0};  // FSAL stage-1 pieces of the current state
    LaneCounters cnt;
    bool drained = false;
    int prio = 0;
    const uint32_t long_ray = prm<BATCH>(Ps, 0).long_ray;
    const int isolate = (int)prm<BATCH>(Ps, 0).isolate;
    Hot H = load_hot(prm<BATCH>(Ps, 0));
    if (lane == 0) atomicMin(&counters->clock_start, (unsigned long long)wall_clock64());

    for (;;) {
        if (L.post) {
            L.post = false;
            post_commit(L, prm<BATCH>(Ps, L.rid), H, cnt);
        }
        if (L.finished) {
            L.finished = false;
            cnt.max_att = max(cnt.max_att, L.ray_att);
            end_sample(L, prm<BATCH>(Ps, L.rid), H);
        }
        // (a wave isolated around a long ray takes no new rays until it ends)
        if (isolate == 0 || prio < isolate)
            refill<BATCH>(L, H, Ps, n_renders, total_items, refill_threshold, counters, lane, drained);
        if (drained && __ballot(L.active) == 0) break;
        // a wave holding a long-running ray (photon-ring orbits) issues
        // ahead of the other wave(s) on its SIMD, so its serial chain of
        // steps, which bounds the frame time, is not slowed by them
        wave_priority(L, long_ray, prio);
        if (!L.active || L.finished) continue;

        // ---- one DP5(4) attempt, split form (see the file comment)
        const G32 g32{(float)H.g.a, (float)H.g.aa, (float)H.g.M, H.g.exact};
        const double aa = H.g.aa;
        const X64 X{L.s.x, L.s.y, L.s.z};
        const X64 Pd{L.s.px, L.s.py, L.s.pz};
        const P32 Pf{(float)Pd.x, (float)Pd.y, (float)Pd.z};
        const float ptf = (float)L.pt;
        if (!L.have_k1) {
            k1m = rhs_mixed(X.x, X.y, X.z, Pf.x, Pf.y, Pf.z, ptf, aa, g32);
            L.have_k1 = true;
            cnt.rhs++;
        }
        const double h = L.h;
        const float hf = (float)h;
        const float hf2 = hf * hf;
        V6 ns;
        K32 esum;     // sum_j dc_j k_j pieces (corr, kp)
        P32 eq;       // sum_k (dc A)_k kp_k
        K32 dsum;     // sum_j d_j k_j pieces (dense output)
        P32 dq;       // sum_k (d A)_k kp_k
        const K32 k1s = k1m;
        {
            const K32 k1 = k1m;
            // stage input from FP64 x + h c p plus the FP32 pieces
#define KT_XIN(c, ci, qsum, asum) (X.c + (h * (ci)) * Pd.c + (double)(hf2 * (qsum) - hf * (asum)))
#define KT_PIN(c, asum) (Pf.c + hf * (asum))
            const K32 k2 = rhs_mixed(KT_XIN(x, KT_C2, 0.0f, F(KT_A21) * k1.cx),
                                     KT_XIN(y, KT_C2, 0.0f, F(KT_A21) * k1.cy),
                                     KT_XIN(z, KT_C2, 0.0f, F(KT_A21) * k1.cz), KT_PIN(x, F(KT_A21) * k1.px),
                                     KT_PIN(y, F(KT_A21) * k1.py), KT_PIN(z, F(KT_A21) * k1.pz), ptf, aa, g32);
#define S3Q(c) (Q31 * k1.c)
#define S3A(c) (F(KT_A31) * k1.c + F(KT_A32) * k2.c)
            const K32 k3 = rhs_mixed(KT_XIN(x, KT_C3, S3Q(px), S3A(cx)), KT_XIN(y, KT_C3, S3Q(py), S3A(cy)),
                                     KT_XIN(z, KT_C3, S3Q(pz), S3A(cz)), KT_PIN(x, S3A(px)), KT_PIN(y, S3A(py)),
                                     KT_PIN(z, S3A(pz)), ptf, aa, g32);
#define S4Q(c) (Q41 * k1.c + Q42 * k2.c)
#define S4A(c) (F(KT_A41) * k1.c + F(KT_A42) * k2.c + F(KT_A43) * k3.c)
            const K32 k4 = rhs_mixed(KT_XIN(x, KT_C4, S4Q(px), S4A(cx)), KT_XIN(y, KT_C4, S4Q(py), S4A(cy)),
                                     KT_XIN(z, KT_C4, S4Q(pz), S4A(cz)), KT_PIN(x, S4A(px)), KT_PIN(y, S4A(py)),
                                     KT_PIN(z, S4A(pz)), ptf, aa, g32);
#define S5Q(c) (Q51 * k1.c + Q52 * k2.c + Q53 * k3.c)
#define S5A(c) (F(KT_A51) * k1.c + F(KT_A52) * k2.c + F(KT_A53) * k3.c + F(KT_A54) * k4.c)
            const K32 k5 = rhs_mixed(KT_XIN(x, KT_C5, S5Q(px), S5A(cx)), KT_XIN(y, KT_C5, S5Q(py), S5A(cy)),
                                     KT_XIN(z, KT_C5, S5Q(pz), S5A(cz)), KT_PIN(x, S5A(px)), KT_PIN(y, S5A(py)),
                                     KT_PIN(z, S5A(pz)), ptf, aa, g32);
#define S6Q(c) (Q61 * k1.c + Q62 * k2.c + Q63 * k3.c + Q64 * k4.c)
#define S6A(c) (F(KT_A61) * k1.c + F(KT_A62) * k2.c + F(KT_A63) * k3.c + F(KT_A64) * k4.c + F(KT_A65) * k5.c)
            const double x6 = KT_XIN(x, 1.0, S6Q(px), S6A(cx)), y6 = KT_XIN(y, 1.0, S6Q(py), S6A(cy));
            const double z6 = KT_XIN(z, 1.0, S6Q(pz), S6A(cz));
            const float px6 = KT_PIN(x, S6A(px)), py6 = KT_PIN(y, S6A(py)), pz6 = KT_PIN(z, S6A(pz));
            // leading terms of the 5th-order and error sums, folded before
            // stage 6 so k2..k5 are dead while stages 6 and 7 run
#define SBQ(c) (BA1 * k1.c + BA2 * k2.c + BA3 * k3.c + BA4 * k4.c + BA5 * k5.c)
#define SBA5(c) (F(KT_B1) * k1.c + F(KT_B3) * k3.c + F(KT_B4) * k4.c + F(KT_B5) * k5.c)
#define SEQ5(c) (EA1 * k1.c + EA2 * k2.c + EA3 * k3.c + EA4 * k4.c + EA5 * k5.c)
#define SEA5(c) (F(KT_E1) * k1.c + F(KT_E3) * k3.c + F(KT_E4) * k4.c + F(KT_E5) * k5.c)
#define SDQ5(c) (DA1 * k1.c + DA2 * k2.c + DA3 * k3.c + DA4 * k4.c + DA5 * k5.c)
#define SDA5(c) (F(DD1) * k1.c + F(DD3) * k3.c + F(DD4) * k4.c + F(DD5) * k5.c)
            dq = P32{SDQ5(px), SDQ5(py), SDQ5(pz)};
            dsum = K32{SDA5(cx), SDA5(cy), SDA5(cz), SDA5(px), SDA5(py), SDA5(pz)};
            const P32 bq{SBQ(px), SBQ(py), SBQ(pz)};
            K32 bsum{SBA5(cx), SBA5(cy), SBA5(cz), SBA5(px), SBA5(py), SBA5(pz)};
            eq = P32{SEQ5(px), SEQ5(py), SEQ5(pz)};
            esum = K32{SEA5(cx), SEA5(cy), SEA5(cz), SEA5(px), SEA5(py), SEA5(pz)};
            const K32 k6 = rhs_mixed(x6, y6, z6, px6, py6, pz6, ptf, aa, g32);
            bsum = K32{fmaf(F(KT_B6), k6.cx, bsum.cx), fmaf(F(KT_B6), k6.cy, bsum.cy), fmaf(F(KT_B6), k6.cz, bsum.cz),
                       fmaf(F(KT_B6), k6.px, bsum.px), fmaf(F(KT_B6), k6.py, bsum.py), fmaf(F(KT_B6), k6.pz, bsum.pz)};
            eq = P32{fmaf(EA6, k6.px, eq.x), fmaf(EA6, k6.py, eq.y), fmaf(EA6, k6.pz, eq.z)};
            dq = P32{fmaf(DA6, k6.px, dq.x), fmaf(DA6, k6.py, dq.y), fmaf(DA6, k6.pz, dq.z)};
            dsum = K32{fmaf(F(DD6), k6.cx, dsum.cx), fmaf(F(DD6), k6.cy, dsum.cy), fmaf(F(DD6), k6.cz, dsum.cz),
                       fmaf(F(DD6), k6.px, dsum.px), fmaf(F(DD6), k6.py, dsum.py), fmaf(F(DD6), k6.pz, dsum.pz)};
            esum = K32{fmaf(F(KT_E6), k6.cx, esum.cx), fmaf(F(KT_E6), k6.cy, esum.cy), fmaf(F(KT_E6), k6.cz, esum.cz),
                       fmaf(F(KT_E6), k6.px, esum.px), fmaf(F(KT_E6), k6.py, esum.py), fmaf(F(KT_E6), k6.pz, esum.pz)};
            // 5th-order solution (FP64 accumulation) = 7th stage input
            const double hh = h * h;
            ns.x = X.x + h * Pd.x + (hh * (double)bq.x - h * (double)bsum.cx);
            ns.y = X.y + h * Pd.y + (hh * (double)bq.y - h * (double)bsum.cy);
            ns.z = X.z + h * Pd.z + (hh * (double)bq.z - h * (double)bsum.cz);
            ns.px = Pd.x + h * (double)bsum.px;
            ns.py = Pd.y + h * (double)bsum.py;
            ns.pz = Pd.z + h * (double)bsum.pz;
#undef KT_XIN
#undef KT_PIN
#undef S3Q
#undef S3A
#undef S4Q
#undef S4A
#undef S5Q
#undef S5A
#undef S6Q
#undef S6A
#undef SBQ
#undef SBA5
#undef SEQ5
#undef SEA5
#undef SDQ5
#undef SDA5
        }
        const K32 k7 = rhs_mixed(ns.x, ns.y, ns.z, (float)ns.px, (float)ns.py, (float)ns.pz, ptf, aa, g32);
        cnt.rhs += 6;
        esum = K32{fmaf(F(KT_E7), k7.cx, esum.cx), fmaf(F(KT_E7), k7.cy, esum.cy), fmaf(F(KT_E7), k7.cz, esum.cz),
                   fmaf(F(KT_E7), k7.px, esum.px), fmaf(F(KT_E7), k7.py, esum.py), fmaf(F(KT_E7), k7.pz, esum.pz)};
        dsum = K32{fmaf(F(DD7), k7.cx, dsum.cx), fmaf(F(DD7), k7.cy, dsum.cy), fmaf(F(DD7), k7.cz, dsum.cz),
                   fmaf(F(DD7), k7.px, dsum.px), fmaf(F(DD7), k7.py, dsum.py), fmaf(F(DD7), k7.pz, dsum.pz)};
        // embedded error: FP32 sums of the gravitational pieces, FP64 norm
        const double tol = L.tol;
        double ex = div_pos(fabs(h * (double)(hf * eq.x - esum.cx)), tol * fmax(fabs(X.x), 1.0));
        double ey = div_pos(fabs(h * (double)(hf * eq.y - esum.cy)), tol * fmax(fabs(X.y), 1.0));
        double ez = div_pos(fabs(h * (double)(hf * eq.z - esum.cz)), tol * fmax(fabs(X.z), 1.0));
        double epx = div_pos(fabs(h * (double)esum.px), tol * fmax(fabs(Pd.x), 1.0));
        double epy = div_pos(fabs(h * (double)esum.py), tol * fmax(fabs(Pd.y), 1.0));
        double epz = div_pos(fabs(h * (double)esum.pz), tol * fmax(fabs(Pd.z), 1.0));
        double err = fmax(fmax(fmax(ex, ey), fmax(ez, epx)), fmax(epy, epz));
        if (__builtin_expect(!(err <= 1.79769313486231570e308), 0)) {
            // a non-finite step error: the library quotients (special values)
            err = fmax(fmax(fmax(fabs(h * (double)(hf * eq.x - esum.cx)) / (tol * fmax(fabs(X.x), 1.0)),
                                 fabs(h * (double)(hf * eq.y - esum.cy)) / (tol * fmax(fabs(X.y), 1.0))),
                            fmax(fabs(h * (double)(hf * eq.z - esum.cz)) / (tol * fmax(fabs(X.z), 1.0)),
                                 fabs(h * (double)esum.px) / (tol * fmax(fabs(Pd.x), 1.0)))),
                       fmax(fabs(h * (double)esum.py) / (tol * fmax(fabs(Pd.y), 1.0)),
                            fabs(h * (double)esum.pz) / (tol * fmax(fabs(Pd.z), 1.0))));
        }
        cnt.attempts++;
        L.ray_att++;
        L.att_in_step++;
        const bool accepted = err <= 1.0;
        L.h = h * 0.9 * pow_m02(err);
        if (!accepted && L.att_in_step < H.max_attempts) continue;

        // ---- commit; the crossing's dense output, root and shading run in
        // FP64 in post_commit from FP64 ingredients rebuilt here:
        // dx/dl_j = p_j - corr_j, sum_j d_j p_j = h sum_k (d A)_k kp_k
        if (accepted) cnt.accepted++;
        if (accepted) k1m = k7;  // FSAL
        L.have_k1 = accepted;
        const V6 s0{X.x, X.y, X.z, Pd.x, Pd.y, Pd.z};
        const V6 ck1{Pd.x - (double)k1s.cx, Pd.y - (double)k1s.cy, Pd.z - (double)k1s.cz,
                     (double)k1s.px, (double)k1s.py, (double)k1s.pz};
        const V6 ck7{ns.px - (double)k7.cx, ns.py - (double)k7.cy, ns.pz - (double)k7.cz,
                     (double)k7.px, (double)k7.py, (double)k7.pz};
        const V6 cd{h * (double)dq.x - (double)dsum.cx, h * (double)dq.y - (double)dsum.cy,
                    h * (double)dq.z - (double)dsum.cz, (double)dsum.px, (double)dsum.py, (double)dsum.pz};
        commit(L, s0, ns, h, ck1, ck7, cd, H, prm<BATCH>(Ps, L.rid).disk_root_mode);
    }

    flush_counters(cnt, counters, lane);
}

hipError_t launch_trace_mixed(const RenderParams* d_params, int n_renders, uint64_t total_items, int refill_threshold,
                              DevCounters* d_counters, int blocks, hipStream_t stream) {
    if (n_renders == 1)
        hipLaunchKernelGGL(trace_mixed_persistent<false>, dim3(blocks), dim3(256), 0, stream, d_params, 1,
                           total_items, refill_threshold, d_counters);
    else
        hipLaunchKernelGGL(trace_mixed_persistent<true>, dim3(blocks), dim3(256), 0, stream, d_params, n_renders,
                           total_items, refill_threshold, d_counters);
    return hipGetLastError();
}

}  // namespace kt
