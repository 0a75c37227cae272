// Host build of the differentiable renderer's adjoint building blocks
// (kerrtrace_amd/csrc/kt_adjoint.hpp) and of the tracer's scalar helpers
// (kt_math.hpp) for CPU checks against the oracle.
// TEST INFRASTRUCTURE: compiled with hipcc --offload-host-only by
// tests/test_adjoint_blocks.py; the product library does not use it.
#include <cstring>
#include <vector>

#include "../../kerrtrace_amd/csrc/kt_adjoint.hpp"

using namespace kt;

static V6 v6(const double* p) { return V6{p[0], p[1], p[2], p[3], p[4], p[5]}; }
static void put(const V6& v, double* p) {
    p[0] = v.x; p[1] = v.y; p[2] = v.z; p[3] = v.px; p[4] = v.py; p[5] = v.pz;
}

extern "C" {

void kta_rhs_vjp(const double Y[6], double pt, double a, double M, const double w[6], double lamY[6],
                 double* lam_pt, double* lam_a) {
    const Geo g = make_geo(a, a * a, M, true);
    V6 l = v6_zero();
    double lp = 0.0, la = 0.0;
    rhs_vjp(v6(Y), pt, g, v6(w), l, lp, la);
    put(l, lamY);
    *lam_pt = lp;
    *lam_a = la;
}

// FP32 variants of the right-hand-side VJP and of one step's reversal.
void kta_rhs_vjp32(const double Y[6], double pt, double a, double M, const double w[6], double lamY[6],
                   double* lam_pt, double* lam_a) {
    V6f l = v6f_zero();
    float lp = 0.f, la = 0.f;
    rhs_vjp32(to_f(v6(Y)), (float)pt, (float)a, (float)(a * a), (float)M, to_f(v6(w)), l, lp, la);
    put(to_d(l), lamY);
    *lam_pt = lp;
    *lam_a = la;
}

void kta_step_reverse32(const double s[6], double h, double pt, double a, double M, const double lam_ns[6],
                        double lam_s[6], double* lam_pt, double* lam_a) {
    const Geo g = make_geo(a, a * a, M, true);
    V6 Y[5];
    StepOut so = dp_step(v6(s), h, pt, g, Y);
    V6f Yf[5];
    for (int i = 0; i < 5; ++i) Yf[i] = to_f(Y[i]);
    float lp = 0.f, la = 0.f;
    V6f ls = dp_step_reverse32(to_f(v6(s)), to_f(so.ns), Yf, (float)h, (float)pt, (float)a, (float)(a * a), (float)M,
                               to_f(v6(lam_ns)), nullptr, v6f_zero(), lp, la);
    put(to_d(ls), lam_s);
    *lam_pt = lp;
    *lam_a = la;
}

void kta_step(const double s[6], double h, double pt, double a, double M, double ns[6]) {
    const Geo g = make_geo(a, a * a, M, true);
    put(dp_step(v6(s), h, pt, g, nullptr).ns, ns);
}

void kta_step_reverse(const double s[6], double h, double pt, double a, double M, const double lam_ns[6],
                      double lam_s[6], double* lam_pt, double* lam_a) {
    const Geo g = make_geo(a, a * a, M, true);
    V6 Y[5];
    StepOut so = dp_step(v6(s), h, pt, g, Y);
    double lp = 0.0, la = 0.0;
    V6 ls = dp_step_reverse(v6(s), so.ns, Y, h, pt, g, v6(lam_ns), nullptr, v6_zero(), lp, la);
    put(ls, lam_s);
    *lam_pt = lp;
    *lam_a = la;
}

// Dense-output crossing from explicit step ingredients s, ns, k1..k7
// (k[6 * i], k2 unused): hit = (t, x, z, px, py, pz) and the adjoints of
// s, ns and k1, k3, k4, k5, k6, k7 for a given adjoint of (x, z, px, py, pz).
void kta_crossing(const double s[6], const double ns[6], const double k[42], double h, int root_mode,
                  const double lam_hit[5], double hit_out[6], double lam_s[6], double lam_ns[6], double lk[36]) {
    const V6 k1 = v6(k), k3 = v6(k + 12), k4 = v6(k + 18), k5 = v6(k + 24), k6 = v6(k + 30), k7 = v6(k + 36);
#define KTA_DS(c) dense_sum(k1.c, k3.c, k4.c, k5.c, k6.c, k7.c)
    const V6 d{KTA_DS(x), KTA_DS(y), KTA_DS(z), KTA_DS(px), KTA_DS(py), KTA_DS(pz)};
#undef KTA_DS
    Crossing c = crossing_hit(v6(s), v6(ns), k1, k7, d, h, root_mode);
    hit_out[0] = c.t; hit_out[1] = c.x; hit_out[2] = c.z; hit_out[3] = c.px; hit_out[4] = c.py; hit_out[5] = c.pz;
    V6 ls = v6_zero(), lns = v6_zero();
    V6 kk[6];
    for (int i = 0; i < 6; ++i) kk[i] = v6_zero();
    crossing_reverse(v6(s), v6(ns), k1, k7, d, h, root_mode, c.t, lam_hit, ls, lns, kk);
    put(ls, lam_s);
    put(lns, lam_ns);
    for (int i = 0; i < 6; ++i) put(kk[i], lk + 6 * i);
}

// Ray start of pixel (px, py) in direction dir (0 spin, 1 inclination,
// 2 distance): vals/tans = (s0[6], pt, kt).
void kta_start(int W, int H, int px, int py, double a, double M, double incl, double r_obs, double fov_scale, int dir,
               double vals[8], double tans[8]) {
    RenderParams P;
    std::memset(&P, 0, sizeof P);
    P.a = a; P.aa = a * a; P.M = M; P.W = W; P.H = H;
    P.inv_w_2 = 2.0 / (double)W; P.inv_h_2 = 2.0 / (double)H; P.aspect = (double)W / (double)H;
    GradConsts G;
    std::memset(&G, 0, sizeof G);
    G.incl = incl; G.r_obs = r_obs; G.fov_scale = fov_scale;
    StartD st = start_d(P, G, px, py, dir);
    for (int k = 0; k < 6; ++k) { vals[k] = st.s[k].v; tans[k] = st.s[k].d; }
    vals[6] = st.pt.v; tans[6] = st.pt.d;
    vals[7] = st.kt.v; tans[7] = st.kt.d;
}

// Emission of a crossing in direction dir (0..4 hit state, 5 pt, 6 kt,
// 7 spin, 8 mdot) with the given LUTs; out = (c0, c1, c2, alpha, lit) values
// and tangents.  sp = {a, M, disk_inner, disk_outer, T_scale, lut_max_temp,
// wl_step, noise_scale, noise_strength, noise_winding, d_rin_da, dTs_da,
// dTs_dm}; ip = {lut_n, noise_detail}.
void kta_shade(const double sp[13], const int ip[2], uint64_t seed_mix, const double* disk_lut, const double* dlut,
               const double* bb_lut, const double hit[5], double pt, double kt, int dir, double vals[5], double tans[4]) {
    RenderParams P;
    std::memset(&P, 0, sizeof P);
    P.a = sp[0]; P.aa = sp[0] * sp[0]; P.M = sp[1]; P.inv_M = 1.0 / sp[1];
    P.disk_inner = sp[2]; P.disk_outer = sp[3]; P.T_scale = sp[4]; P.lut_max_temp = sp[5]; P.wl_step = sp[6];
    P.noise_scale = sp[7]; P.noise_strength = sp[8]; P.noise_winding = sp[9]; P.noise_seed_mix = seed_mix;
    P.lut_n = ip[0]; P.noise_detail = ip[1];
    P.disk_lut = disk_lut; P.bb_lut = bb_lut;
    GradConsts G;
    std::memset(&G, 0, sizeof G);
    G.dlut = dlut; G.d_rin_da = sp[10]; G.dTs_da = sp[11]; G.dTs_dm = sp[12];
    const double v[7] = {hit[0], hit[1], hit[2], hit[3], hit[4], pt, kt};
    D1 in[7];
    for (int k = 0; k < 7; ++k) in[k] = D1{v[k], dir == k ? 1.0 : 0.0};
    const bool da = dir == 7, dm = dir == 8;
    ShadeD o = shade_d(P, G, in[0], in[1], in[2], in[3], in[4], in[5], in[6], D1{P.a, da ? 1.0 : 0.0},
                       D1{P.aa, da ? 2.0 * P.a : 0.0}, D1{P.disk_inner, da ? G.d_rin_da : 0.0},
                       D1{P.T_scale, da ? G.dTs_da : (dm ? G.dTs_dm : 0.0)}, da ? 1.0 : 0.0);
    for (int k = 0; k < 3; ++k) { vals[k] = o.c[k].v; tans[k] = o.c[k].d; }
    vals[3] = o.alpha.v; tans[3] = o.alpha.d;
    vals[4] = o.lit ? 1.0 : 0.0;
}

// the FP32 stage inputs of a step (dp_stages32) next to the FP64 ones of dp_step
void kta_stages32(const double s[6], double h, double pt, double a, double M, double y64[30], double y32[30]) {
    const Geo g = make_geo(a, a * a, M, true);
    V6 Y[5];
    (void)dp_step(v6(s), h, pt, g, Y);
    V6f Yf[5];
    dp_stages32(to_f(v6(s)), (float)h, (float)pt, (float)a, (float)(a * a), (float)M, Yf);
    for (int i = 0; i < 5; ++i) {
        put(Y[i], y64 + 6 * i);
        put(to_d(Yf[i]), y32 + 6 * i);
    }
}

// the tracer's step-size controller factor x^(-1/5) (kt_math.hpp pow_m02)
double kta_pow_m02(double x) { return pow_m02(x); }

}  // extern "C"
