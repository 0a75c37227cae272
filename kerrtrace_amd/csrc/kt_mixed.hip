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
// controller, the dense-output disk root (the crossing step's stage
// derivatives are rebuilt in FP64 from the FP32 pieces) and shading.
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

struct X64 {
    double x, y, z;
};
struct P32 {
    float x, y, z;
};

template <bool BATCH>
__global__ __launch_bounds__(256) void trace_mixed_persistent(const RenderParams* __restrict__ Ps, int n_renders,
                                                              uint64_t total_items, int refill_threshold,
                                                              DevCounters* __restrict__ counters) {
    const int lane = (int)(threadIdx.x & (kWave - 1));
    Lane L;
    L.active = false;
    L.rid = 0;
    L.overflow = false;
    L.overflow_T = 0.0;
    L.ray_att = 0;
    K32 k1m{0, 0, 0, 0, 0, 0};  // FSAL stage-1 pieces of the current state
    LaneCounters cnt;
    bool drained = false;
    Hot H = load_hot(prm<BATCH>(Ps, 0));
    G32 g32{(float)H.g.a, (float)H.g.aa, (float)H.g.M, H.g.exact};

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
                        if (BATCH) {
                            H = load_hot(P);
                            g32 = G32{(float)H.g.a, (float)H.g.aa, (float)H.g.M, H.g.exact};
                        }
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

        // ---- one DP5(4) attempt, split form (see the file comment)
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
        const K32 k1 = k1m;
        const double h = L.h;
        const float hf = (float)h;
        const float hf2 = hf * hf;
        // stage input from FP64 x + h c p plus the FP32 pieces
#define KT_XIN(c, ci, qsum, asum) (X.c + (h * (ci)) * Pd.c + (double)(hf2 * (qsum) - hf * (asum)))
#define KT_PIN(c, asum) (Pf.c + hf * (asum))
        // stage 2
        K32 k2 = rhs_mixed(KT_XIN(x, KT_C2, 0.0f, F(KT_A21) * k1.cx), KT_XIN(y, KT_C2, 0.0f, F(KT_A21) * k1.cy),
                           KT_XIN(z, KT_C2, 0.0f, F(KT_A21) * k1.cz), KT_PIN(x, F(KT_A21) * k1.px),
                           KT_PIN(y, F(KT_A21) * k1.py), KT_PIN(z, F(KT_A21) * k1.pz), ptf, aa, g32);
        // stage 3
#define S3Q(c) (Q31 * k1.c)
#define S3A(c, d) (F(KT_A31) * k1.c + F(KT_A32) * k2.c)
        K32 k3 = rhs_mixed(KT_XIN(x, KT_C3, S3Q(px), S3A(cx, 0)), KT_XIN(y, KT_C3, S3Q(py), S3A(cy, 0)),
                           KT_XIN(z, KT_C3, S3Q(pz), S3A(cz, 0)), KT_PIN(x, S3A(px, 0)), KT_PIN(y, S3A(py, 0)),
                           KT_PIN(z, S3A(pz, 0)), ptf, aa, g32);
        // stage 4
#define S4Q(c) (Q41 * k1.c + Q42 * k2.c)
#define S4A(c) (F(KT_A41) * k1.c + F(KT_A42) * k2.c + F(KT_A43) * k3.c)
        K32 k4 = rhs_mixed(KT_XIN(x, KT_C4, S4Q(px), S4A(cx)), KT_XIN(y, KT_C4, S4Q(py), S4A(cy)),
                           KT_XIN(z, KT_C4, S4Q(pz), S4A(cz)), KT_PIN(x, S4A(px)), KT_PIN(y, S4A(py)),
                           KT_PIN(z, S4A(pz)), ptf, aa, g32);
        // stage 5
#define S5Q(c) (Q51 * k1.c + Q52 * k2.c + Q53 * k3.c)
#define S5A(c) (F(KT_A51) * k1.c + F(KT_A52) * k2.c + F(KT_A53) * k3.c + F(KT_A54) * k4.c)
        K32 k5 = rhs_mixed(KT_XIN(x, KT_C5, S5Q(px), S5A(cx)), KT_XIN(y, KT_C5, S5Q(py), S5A(cy)),
                           KT_XIN(z, KT_C5, S5Q(pz), S5A(cz)), KT_PIN(x, S5A(px)), KT_PIN(y, S5A(py)),
                           KT_PIN(z, S5A(pz)), ptf, aa, g32);
        // stage 6 (c6 = 1)
#define S6Q(c) (Q61 * k1.c + Q62 * k2.c + Q63 * k3.c + Q64 * k4.c)
#define S6A(c) (F(KT_A61) * k1.c + F(KT_A62) * k2.c + F(KT_A63) * k3.c + F(KT_A64) * k4.c + F(KT_A65) * k5.c)
        K32 k6 = rhs_mixed(KT_XIN(x, 1.0, S6Q(px), S6A(cx)), KT_XIN(y, 1.0, S6Q(py), S6A(cy)),
                           KT_XIN(z, 1.0, S6Q(pz), S6A(cz)), KT_PIN(x, S6A(px)), KT_PIN(y, S6A(py)),
                           KT_PIN(z, S6A(pz)), ptf, aa, g32);
        // 5th-order solution (FP64 accumulation) = 7th stage input
#define SBQ(c) (BA1 * k1.c + BA2 * k2.c + BA3 * k3.c + BA4 * k4.c + BA5 * k5.c)
#define SBA(c) (F(KT_B1) * k1.c + F(KT_B3) * k3.c + F(KT_B4) * k4.c + F(KT_B5) * k5.c + F(KT_B6) * k6.c)
        const double hh = h * h;
        V6 ns;
        ns.x = X.x + h * Pd.x + (hh * (double)SBQ(px) - h * (double)SBA(cx));
        ns.y = X.y + h * Pd.y + (hh * (double)SBQ(py) - h * (double)SBA(cy));
        ns.z = X.z + h * Pd.z + (hh * (double)SBQ(pz) - h * (double)SBA(cz));
        ns.px = Pd.x + h * (double)SBA(px);
        ns.py = Pd.y + h * (double)SBA(py);
        ns.pz = Pd.z + h * (double)SBA(pz);
        K32 k7 = rhs_mixed(ns.x, ns.y, ns.z, (float)ns.px, (float)ns.py, (float)ns.pz, ptf, aa, g32);
        cnt.rhs += 6;
        // embedded error: FP32 sums of the gravitational pieces, FP64 norm
#define SEQ(c) (EA1 * k1.c + EA2 * k2.c + EA3 * k3.c + EA4 * k4.c + EA5 * k5.c + EA6 * k6.c)
#define SEA(c) (F(KT_E1) * k1.c + F(KT_E3) * k3.c + F(KT_E4) * k4.c + F(KT_E5) * k5.c + F(KT_E6) * k6.c + F(KT_E7) * k7.c)
        const double tol = L.tol;
        double ex = fabs(h * (double)(hf * SEQ(px) - SEA(cx))) / (tol * fmax(fabs(X.x), 1.0));
        double ey = fabs(h * (double)(hf * SEQ(py) - SEA(cy))) / (tol * fmax(fabs(X.y), 1.0));
        double ez = fabs(h * (double)(hf * SEQ(pz) - SEA(cz))) / (tol * fmax(fabs(X.z), 1.0));
        double epx = fabs(h * (double)SEA(px)) / (tol * fmax(fabs(Pd.x), 1.0));
        double epy = fabs(h * (double)SEA(py)) / (tol * fmax(fabs(Pd.y), 1.0));
        double epz = fabs(h * (double)SEA(pz)) / (tol * fmax(fabs(Pd.z), 1.0));
        double err = fmax(fmax(fmax(ex, ey), fmax(ez, epx)), fmax(epy, epz));
        cnt.attempts++;
        L.ray_att++;
        L.att_in_step++;
        const bool accepted = err <= 1.0;
        L.h = h * 0.9 * pow_m02(err);
        if (!accepted && L.att_in_step < H.max_attempts) continue;

        // ---- commit
        if (accepted) cnt.accepted++;
        const V6 s0 = L.s;
        L.s = ns;
        L.att_in_step = 0;
        double r_now = ks_r(ns.x, ns.y, ns.z, aa);
        bool done = false;
        const RenderParams& P = prm<BATCH>(Ps, L.rid);
        if (r_now < H.rh_eps) {
            L.cls = (L.cls & ~3) | 1;
            done = true;
        } else {
            if ((s0.y < 0.0 && ns.y >= 0.0) || (s0.y > 0.0 && ns.y <= 0.0)) {
                cnt.crossings++;
                // FP64 stage derivatives of this step: dx/dl_j = p_j - corr_j
#define KX(c, pc, j, asum) ((Pd.pc + h * (double)(asum)) - (double)j.c)
                const double k1y = Pd.y - (double)k1.cy;
                const double k3y = KX(cy, y, k3, S3A(py, 0));
                const double k4y = KX(cy, y, k4, S4A(py));
                const double k5y = KX(cy, y, k5, S5A(py));
                const double k6y = KX(cy, y, k6, S6A(py));
                const double k7y = ns.py - (double)k7.cy;
                double t;
                if (P.disk_root_mode == 1) {
                    t = disk_root(s0.y, ns.y, k1y, k3y, k4y, k5y, k6y, k7y, h);
                } else {
                    double den = s0.y - ns.y;
                    t = den != 0.0 ? s0.y / den : 0.5;
                    t = fmin(fmax(t, 0.0), 1.0);
                }
                double x_hit = dense(s0.x, ns.x, Pd.x - (double)k1.cx, KX(cx, x, k3, S3A(px, 0)),
                                     KX(cx, x, k4, S4A(px)), KX(cx, x, k5, S5A(px)), KX(cx, x, k6, S6A(px)),
                                     ns.px - (double)k7.cx, h, t);
                double z_hit = dense(s0.z, ns.z, Pd.z - (double)k1.cz, KX(cz, z, k3, S3A(pz, 0)),
                                     KX(cz, z, k4, S4A(pz)), KX(cz, z, k5, S5A(pz)), KX(cz, z, k6, S6A(pz)),
                                     ns.pz - (double)k7.cz, h, t);
                double px_hit = dense(s0.px, ns.px, (double)k1.px, (double)k3.px, (double)k4.px, (double)k5.px,
                                      (double)k6.px, (double)k7.px, h, t);
                double py_hit = dense(s0.py, ns.py, (double)k1.py, (double)k3.py, (double)k4.py, (double)k5.py,
                                      (double)k6.py, (double)k7.py, h, t);
                double pz_hit = dense(s0.pz, ns.pz, (double)k1.pz, (double)k3.pz, (double)k4.pz, (double)k5.pz,
                                      (double)k6.pz, (double)k7.pz, h, t);
#undef KX
                if (shade_crossing(P, x_hit, z_hit, px_hit, py_hit, pz_hit, L.pt, L.kt, L.col, L.Tr, L.overflow,
                                   L.overflow_T)) {
                    L.cls |= 4;
                    if (((L.cls >> 5) & 7) < 7) L.cls += 32;  // contributions, saturating at 7
                }
            }
            if (r_now > H.R_esc) {
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
                done = true;
            }
        }
        if (accepted) k1m = k7;  // FSAL
        L.have_k1 = accepted;
        if (!done) {
            L.step += 1;
            if (L.step < H.max_steps && L.Tr > H.cutoff) {
                L.tol = step_tolerance(L.s, r_now, H.g, H.tol, H.tol_floor);
            } else {
                done = true;
            }
        }
        if (done) {
            cnt.max_att = max(cnt.max_att, L.ray_att);
            end_sample(L, P, H);
        }
    }
#undef KT_XIN
#undef KT_PIN

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
