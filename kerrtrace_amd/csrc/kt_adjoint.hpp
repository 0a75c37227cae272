// Reverse-mode building blocks of the differentiable renderer (host +
// device, so a host build of the same source can be checked on the CPU).
//
// Differentiated map.  One ray of the observer scene as a function of
// theta = (spin a, inclination i [rad], observer distance r_obs, accretion
// scale mdot) with its committed step sequence frozen: the adaptive
// controller chooses the steps, but the derivative is that of the
// composition of the DP5(4) updates s_{n+1} = s_n + h_n sum_j b_j k_j with
// those h_n held fixed (discretise-then-differentiate).  Event decisions
// (crossings, horizon, escape, LUT cell) are held fixed as well.
//
// Right-hand side adjoint.  With the exact right-hand side f is the
// symplectic gradient of the Kerr-Schild Hamiltonian,
//   f = J grad Ham,  J (u, w) = (w, -u),
//   Ham(x, p) = p.p / 2 - H(x) (l(x).p - pt)^2   (+ a constant in pt),
// so the Jacobian is J Hess, and because Hess is symmetric the
// vector-Jacobian product w^T df/dz = Hess J^T w = J^T (J Hess)(J^T w) is
// one Jacobian-vector product along v = J^T w = (-w_p, w_x): one pass of a
// forward-mode dual number through the right-hand side.  The parameter
// adjoints come from the same pass:
//   w . df/dpt = D_v (dHam/dpt) = D_v (2 H lp),
//   w . df/da  = D_v (dHam/da) = -D_v (H_a lp^2 + 2 H lp (l_a . p)),
// with the a-partials of r, l and H at fixed position written out below.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"
#include "kt_math.hpp"

namespace kt {

// ------------------------------------------------------------ dual number
struct D1 {
    double v, d;
};
KT_HDI D1 cst(double v) { return D1{v, 0.0}; }
KT_HDI D1 operator+(D1 a, D1 b) { return D1{a.v + b.v, a.d + b.d}; }
KT_HDI D1 operator-(D1 a, D1 b) { return D1{a.v - b.v, a.d - b.d}; }
KT_HDI D1 operator-(D1 a) { return D1{-a.v, -a.d}; }
KT_HDI D1 operator*(D1 a, D1 b) { return D1{a.v * b.v, fma(a.d, b.v, a.v * b.d)}; }
KT_HDI D1 operator*(double a, D1 b) { return D1{a * b.v, a * b.d}; }
KT_HDI D1 operator*(D1 a, double b) { return D1{a.v * b, a.d * b}; }
KT_HDI D1 operator+(D1 a, double b) { return D1{a.v + b, a.d}; }
KT_HDI D1 operator+(double a, D1 b) { return D1{a + b.v, b.d}; }
KT_HDI D1 operator-(D1 a, double b) { return D1{a.v - b, a.d}; }
KT_HDI D1 operator-(double a, D1 b) { return D1{a - b.v, -b.d}; }
KT_HDI D1 operator/(D1 a, D1 b) {
    double q = a.v / b.v;
    return D1{q, (a.d - q * b.d) / b.v};
}
KT_HDI D1 operator/(double a, D1 b) {
    double q = a / b.v;
    return D1{q, -(q * b.d) / b.v};
}
KT_HDI D1 operator/(D1 a, double b) { return D1{a.v / b, a.d / b}; }
KT_HDI D1 dfma(D1 a, D1 b, D1 c) { return D1{fma(a.v, b.v, c.v), fma(a.d, b.v, fma(a.v, b.d, c.d))}; }
KT_HDI D1 dsqrt(D1 a) {
    double r = sqrt(a.v);
    return D1{r, a.d * (0.5 / r)};
}
KT_HDI D1 dexp(D1 a) {
    double e = exp(a.v);
    return D1{e, e * a.d};
}
KT_HDI void dsincos(D1 a, D1& s, D1& c) {
    double sv, cv;
    sincos(a.v, &sv, &cv);
    s = D1{sv, cv * a.d};
    c = D1{cv, -sv * a.d};
}
KT_HDI D1 datan2(D1 y, D1 x) {
    double q = fma(x.v, x.v, y.v * y.v);
    return D1{atan2(y.v, x.v), (x.v * y.d - y.v * x.d) / q};
}

KT_HDI V6 v6_zero() { return V6{0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; }
KT_HDI void v6_axpy(V6& acc, double a, const V6& x) {
    acc.x = fma(a, x.x, acc.x);
    acc.y = fma(a, x.y, acc.y);
    acc.z = fma(a, x.z, acc.z);
    acc.px = fma(a, x.px, acc.px);
    acc.py = fma(a, x.py, acc.py);
    acc.pz = fma(a, x.pz, acc.pz);
}
KT_HDI void v6_add(V6& acc, const V6& x) {
    acc.x += x.x;
    acc.y += x.y;
    acc.z += x.z;
    acc.px += x.px;
    acc.py += x.py;
    acc.pz += x.pz;
}

// --------------------------------------------------- right-hand side VJP
// lamY += w^T df/dY at stage input Y; lam_pt += w . df/dpt;
// lam_a += w . df/da (exact right-hand side only).  The dual pass runs the
// factored right-hand side of rhs() (one reciprocal for 1/s, 1/r,
// 1/(r^2+a^2), 1/(r^4+a^2y^2)); only its tangents are used.
KT_HDI D1 drcp(D1 b) {
    const double iv = div_pos(1.0, b.v);
    return D1{iv, -(b.d * iv) * iv};
}
KT_HDI D1 dsqrt_pos(D1 a) {
    const double r = sqrt_pos(a.v);
    return D1{r, a.d * (0.5 * div_pos(1.0, r))};
}
KT_HDI void rhs_vjp(const V6& Y, double pt, const Geo& g, const V6& w, V6& lamY, double& lam_pt, double& lam_a) {
    const D1 x{Y.x, -w.px}, y{Y.y, -w.py}, z{Y.z, -w.pz};
    const D1 px{Y.px, w.x}, py{Y.py, w.y}, pz{Y.pz, w.z};
    const double a2 = g.aa, a = g.a, M = g.M;
    const D1 rho2 = dfma(x, x, dfma(y, y, z * z));
    const D1 u = rho2 - a2;
    const D1 ya = a2 * y;
    const D1 yy = y * y;
    const D1 s = dsqrt_pos(dfma(u, u, 4.0 * (ya * y)));
    const D1 r2 = 0.5 * (u + s);
    const D1 r = dsqrt_pos(r2);
    const D1 denom = r2 + a2;
    const D1 denomH = dfma(r2, r2, ya * y);
    const D1 sr = s * r;
    const D1 dd = denom * denomH;
    const D1 q = drcp(sr * dd);
    const D1 inv_s = (r * dd) * q;
    const D1 inv_r = (s * dd) * q;
    const D1 inv_denom = (sr * denomH) * q;
    const D1 inv_denomH = (sr * denom) * q;
    const D1 Nx = dfma(r, x, -(a * z));
    const D1 Nz = dfma(r, z, a * x);
    const D1 lx = Nx * inv_denom, ly = y * inv_r, lz = Nz * inv_denom;
    const D1 r3 = r2 * r;
    const D1 H = (M * r3) * inv_denomH;
    const D1 Q = dfma(lx, px, lz * pz);
    const D1 lp = dfma(ly, py, Q) - pt;
    const D1 c = (2.0 * H) * lp;
    const D1 fx = px - c * lx, fy = py - c * ly, fz = pz - c * lz;
    const D1 inv_2r = 0.5 * inv_r;
    const D1 w1 = (u * inv_s + 1.0) * inv_2r;
    const D1 w2 = ((u + 2.0 * a2) * inv_s + 1.0) * inv_2r;
    const D1 XP = dfma(x, px, z * pz);
    const D1 inv_r2 = inv_r * inv_r;
    const D1 G = dfma(-2.0 * r, Q, XP) * inv_denom - (y * py) * inv_r2;
    const D1 inv_denomH2 = inv_denomH * inv_denomH;
    const D1 HR = ((M * r2) * inv_denomH2) * (3.0 * (ya * y) - r2 * r2);
    const D1 HY = -(2.0 * M) * ya * (r3 * inv_denomH2);
    const D1 lp2 = lp * lp;
    const D1 Z = dfma(HR, lp2, c * G);
    const D1 cD = c * inv_denom;
    const D1 fpx = dfma(x * w1, Z, cD * dfma(r, px, a * pz));
    const D1 fpy = dfma(y * w2, Z, dfma(HY, lp2, (c * py) * inv_r));
    const D1 fpz = dfma(z * w1, Z, cD * dfma(r, pz, -(a * px)));
    // w^T df/dY = J^T (tangent of f along v)
    lamY.x -= fpx.d;
    lamY.y -= fpy.d;
    lamY.z -= fpz.d;
    lamY.px += fx.d;
    lamY.py += fy.d;
    lamY.pz += fz.d;
    // parameters: pt enters Ham through lp; a through r, l and H
    lam_pt += 2.0 * fma(H.d, lp.v, H.v * lp.d);
    const D1 r2_a = a * ((2.0 * yy - u) * inv_s - 1.0);
    const D1 r_a = r2_a * inv_2r;
    const D1 D_a = r2_a + 2.0 * a;
    const D1 inv_denom2 = inv_denom * inv_denom;
    const D1 lx_a = ((r_a * x - z) * denom - Nx * D_a) * inv_denom2;
    const D1 lz_a = ((r_a * z + x) * denom - Nz * D_a) * inv_denom2;
    const D1 ly_a = -(y * r_a) * inv_r2;
    const D1 DH_a = 2.0 * r2 * r2_a + (2.0 * a) * yy;
    const D1 H_a = (M * inv_denomH2) * ((3.0 * r2) * denomH * r_a - r3 * DH_a);
    const D1 lp_a = dfma(lx_a, px, dfma(ly_a, py, lz_a * pz));
    const D1 ham_a = -(H_a * lp2 + 2.0 * H * lp * lp_a);
    lam_a += ham_a.d;
}

// ------------------------------------------------------- DP5(4) stepping
// The committed update of the tracer with its arithmetic: stage inputs as
// fma chains, the 5th-order solution folded as the tracer folds it, and the
// dense-output sum d1 k1 + d3 k3 + ... + d6 k6 left to right (d7 k7 is
// added by the caller when it needs k7).  Y receives the inputs of stages
// 2..6 when non-null.
struct StepOut {
    V6 ns, k1, dacc6;
};
#define KTA_A21 (1.0 / 5.0)
#define KTA_A31 (3.0 / 40.0)
#define KTA_A32 (9.0 / 40.0)
#define KTA_A41 (44.0 / 45.0)
#define KTA_A42 (-56.0 / 15.0)
#define KTA_A43 (32.0 / 9.0)
#define KTA_A51 (19372.0 / 6561.0)
#define KTA_A52 (-25360.0 / 2187.0)
#define KTA_A53 (64448.0 / 6561.0)
#define KTA_A54 (-212.0 / 729.0)
#define KTA_A61 (9017.0 / 3168.0)
#define KTA_A62 (-355.0 / 33.0)
#define KTA_A63 (46732.0 / 5247.0)
#define KTA_A64 (49.0 / 176.0)
#define KTA_A65 (-5103.0 / 18656.0)
#define KTA_B1 (35.0 / 384.0)
#define KTA_B3 (500.0 / 1113.0)
#define KTA_B4 (125.0 / 192.0)
#define KTA_B5 (-2187.0 / 6784.0)
#define KTA_B6 (11.0 / 84.0)

// The forward pass's right-hand side: rhs() (FP64 tracer) or rhs32 (mixed
// tracer, device only).
template <bool MIXED>
KT_HDI V6 rhs_fwd(const V6& st, double pt, const Geo& g) {
#if defined(__HIP_DEVICE_COMPILE__)
    if (MIXED) return rhs32(st, (float)pt, g, make_g32(g));
#endif
    return rhs(st, pt, g);
}

template <bool MIXED = false>
KT_HDI StepOut dp_step(const V6& s, double hs, double pt, const Geo& g, V6* Y) {
#define KTA_ST2(c) fma(hs, KTA_A21 * k1.c, s.c)
#define KTA_ST3(c) fma(hs, fma(KTA_A32, k2.c, KTA_A31 * k1.c), s.c)
#define KTA_ST4(c) fma(hs, fma(KTA_A43, k3.c, fma(KTA_A42, k2.c, KTA_A41 * k1.c)), s.c)
#define KTA_ST5(c) fma(hs, fma(KTA_A54, k4.c, fma(KTA_A53, k3.c, fma(KTA_A52, k2.c, KTA_A51 * k1.c))), s.c)
#define KTA_ST6(c) \
    fma(hs, fma(KTA_A65, k5.c, fma(KTA_A64, k4.c, fma(KTA_A63, k3.c, fma(KTA_A62, k2.c, KTA_A61 * k1.c)))), s.c)
#define KTA_V(M) V6{M(x), M(y), M(z), M(px), M(py), M(pz)}
    StepOut o;
    const V6 k1 = rhs_fwd<MIXED>(s, pt, g);
    const V6 y2 = KTA_V(KTA_ST2);
    const V6 k2 = rhs_fwd<MIXED>(y2, pt, g);
    const V6 y3 = KTA_V(KTA_ST3);
    const V6 k3 = rhs_fwd<MIXED>(y3, pt, g);
    const V6 y4 = KTA_V(KTA_ST4);
    const V6 k4 = rhs_fwd<MIXED>(y4, pt, g);
    const V6 y5 = KTA_V(KTA_ST5);
    const V6 k5 = rhs_fwd<MIXED>(y5, pt, g);
    const V6 y6 = KTA_V(KTA_ST6);
#define KTA_B4ACC(c) fma(KTA_B5, k5.c, fma(KTA_B4, k4.c, fma(KTA_B3, k3.c, KTA_B1 * k1.c)))
#define KTA_D5ACC(c) fma(Dense::d5, k5.c, fma(Dense::d4, k4.c, fma(Dense::d3, k3.c, Dense::d1 * k1.c)))
    const V6 bacc = KTA_V(KTA_B4ACC);
    const V6 dacc = KTA_V(KTA_D5ACC);
    const V6 k6 = rhs_fwd<MIXED>(y6, pt, g);
#define KTA_NS(c) fma(hs, fma(KTA_B6, k6.c, bacc.c), s.c)
#define KTA_D6ACC(c) fma(Dense::d6, k6.c, dacc.c)
    o.ns = KTA_V(KTA_NS);
    o.dacc6 = KTA_V(KTA_D6ACC);
    o.k1 = k1;
    if (Y) {
        Y[0] = y2;
        Y[1] = y3;
        Y[2] = y4;
        Y[3] = y5;
        Y[4] = y6;
    }
    return o;
#undef KTA_ST2
#undef KTA_ST3
#undef KTA_ST4
#undef KTA_ST5
#undef KTA_ST6
#undef KTA_B4ACC
#undef KTA_D5ACC
#undef KTA_NS
#undef KTA_D6ACC
}

// Reverse of one committed step.  Y[0..4] are the inputs of stages 2..6;
// lam_ns is the adjoint of the step's end state; lk[0..5] are direct
// adjoints of k1, k3, k4, k5, k6, k7 (the crossing's dense output), and
// lam_s_dir the direct adjoint of the start state.  Returns the adjoint
// of the start state and accumulates the pt and spin adjoints.
KT_HDI V6 dp_step_reverse(const V6& s, const V6& ns, const V6* Y, double hs, double pt, const Geo& g, V6 lam_ns,
                          const V6* lk, const V6& lam_s_dir, double& lam_pt, double& lam_a) {
    if (lk) rhs_vjp(ns, pt, g, lk[5], lam_ns, lam_pt, lam_a);  // k7 = f(ns)
    V6 l1 = v6_zero(), l2 = v6_zero(), l3 = v6_zero(), l4 = v6_zero(), l5 = v6_zero(), l6 = v6_zero();
    v6_axpy(l1, hs * KTA_B1, lam_ns);
    v6_axpy(l3, hs * KTA_B3, lam_ns);
    v6_axpy(l4, hs * KTA_B4, lam_ns);
    v6_axpy(l5, hs * KTA_B5, lam_ns);
    v6_axpy(l6, hs * KTA_B6, lam_ns);
    if (lk) {
        v6_add(l1, lk[0]);
        v6_add(l3, lk[1]);
        v6_add(l4, lk[2]);
        v6_add(l5, lk[3]);
        v6_add(l6, lk[4]);
    }
    V6 ls = lam_ns;
    v6_add(ls, lam_s_dir);
    V6 lY;
    // stage 6: Y6 = s + h (a61 k1 + ... + a65 k5)
    lY = v6_zero();
    rhs_vjp(Y[4], pt, g, l6, lY, lam_pt, lam_a);
    v6_add(ls, lY);
    v6_axpy(l1, hs * KTA_A61, lY);
    v6_axpy(l2, hs * KTA_A62, lY);
    v6_axpy(l3, hs * KTA_A63, lY);
    v6_axpy(l4, hs * KTA_A64, lY);
    v6_axpy(l5, hs * KTA_A65, lY);
    // stage 5
    lY = v6_zero();
    rhs_vjp(Y[3], pt, g, l5, lY, lam_pt, lam_a);
    v6_add(ls, lY);
    v6_axpy(l1, hs * KTA_A51, lY);
    v6_axpy(l2, hs * KTA_A52, lY);
    v6_axpy(l3, hs * KTA_A53, lY);
    v6_axpy(l4, hs * KTA_A54, lY);
    // stage 4
    lY = v6_zero();
    rhs_vjp(Y[2], pt, g, l4, lY, lam_pt, lam_a);
    v6_add(ls, lY);
    v6_axpy(l1, hs * KTA_A41, lY);
    v6_axpy(l2, hs * KTA_A42, lY);
    v6_axpy(l3, hs * KTA_A43, lY);
    // stage 3
    lY = v6_zero();
    rhs_vjp(Y[1], pt, g, l3, lY, lam_pt, lam_a);
    v6_add(ls, lY);
    v6_axpy(l1, hs * KTA_A31, lY);
    v6_axpy(l2, hs * KTA_A32, lY);
    // stage 2
    lY = v6_zero();
    rhs_vjp(Y[0], pt, g, l2, lY, lam_pt, lam_a);
    v6_add(ls, lY);
    v6_axpy(l1, hs * KTA_A21, lY);
    // stage 1 at s
    rhs_vjp(s, pt, g, l1, ls, lam_pt, lam_a);
    return ls;
}

// ------------------------------------------------ FP32 adjoint variant
// The step reversal with its right-hand-side VJPs, stage adjoints and state
// adjoint in FP32 (the forward states and stage inputs come from the FP64
// re-integration and are rounded once): half the registers of the FP64
// reversal.  Only the VJPs' evaluation points and the adjoint recursion
// lose digits; the crossing adjoint (dense-output root, emission) stays FP64.
struct V6f {
    float x, y, z, px, py, pz;
};
KT_HDI V6f v6f_zero() { return V6f{0.f, 0.f, 0.f, 0.f, 0.f, 0.f}; }
KT_HDI V6f to_f(const V6& v) {
    return V6f{(float)v.x, (float)v.y, (float)v.z, (float)v.px, (float)v.py, (float)v.pz};
}
KT_HDI V6 to_d(const V6f& v) {
    return V6{(double)v.x, (double)v.y, (double)v.z, (double)v.px, (double)v.py, (double)v.pz};
}
KT_HDI void v6f_axpy(V6f& acc, float a, const V6f& x) {
    acc.x = fmaf(a, x.x, acc.x);
    acc.y = fmaf(a, x.y, acc.y);
    acc.z = fmaf(a, x.z, acc.z);
    acc.px = fmaf(a, x.px, acc.px);
    acc.py = fmaf(a, x.py, acc.py);
    acc.pz = fmaf(a, x.pz, acc.pz);
}
KT_HDI void v6f_add(V6f& acc, const V6f& x) {
    acc.x += x.x;
    acc.y += x.y;
    acc.z += x.z;
    acc.px += x.px;
    acc.py += x.py;
    acc.pz += x.pz;
}
// FP32 dual number
struct F1 {
    float v, d;
};
KT_HDI F1 operator+(F1 a, F1 b) { return F1{a.v + b.v, a.d + b.d}; }
KT_HDI F1 operator-(F1 a, F1 b) { return F1{a.v - b.v, a.d - b.d}; }
KT_HDI F1 operator-(F1 a) { return F1{-a.v, -a.d}; }
KT_HDI F1 operator*(F1 a, F1 b) { return F1{a.v * b.v, fmaf(a.d, b.v, a.v * b.d)}; }
KT_HDI F1 operator*(float a, F1 b) { return F1{a * b.v, a * b.d}; }
KT_HDI F1 operator+(F1 a, float b) { return F1{a.v + b, a.d}; }
KT_HDI F1 operator-(F1 a, float b) { return F1{a.v - b, a.d}; }
KT_HDI F1 ffma(F1 a, F1 b, F1 c) { return F1{fmaf(a.v, b.v, c.v), fmaf(a.d, b.v, fmaf(a.v, b.d, c.d))}; }
KT_HDI float rcp32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
KT_HDI float rsq32(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}
KT_HDI F1 frcp(F1 b) {
    const float iv = rcp32(b.v);
    return F1{iv, -(b.d * iv) * iv};
}
KT_HDI F1 fsqrt(F1 a) {
    const float rs = rsq32(a.v);
    return F1{a.v * rs, a.d * (0.5f * rs)};
}

// rhs_vjp in FP32 (the same dual pass; see rhs_vjp).
KT_HDI void rhs_vjp32(const V6f& Y, float pt, float a, float a2, float M, const V6f& w, V6f& lamY, float& lam_pt,
                      float& lam_a) {
    const F1 x{Y.x, -w.px}, y{Y.y, -w.py}, z{Y.z, -w.pz};
    const F1 px{Y.px, w.x}, py{Y.py, w.y}, pz{Y.pz, w.z};
    const F1 rho2 = ffma(x, x, ffma(y, y, z * z));
    const F1 u = rho2 - a2;
    const F1 ya = a2 * y;
    const F1 yy = y * y;
    const F1 s = fsqrt(ffma(u, u, 4.0f * (ya * y)));
    const F1 r2 = 0.5f * (u + s);
    const F1 r = fsqrt(r2);
    const F1 denom = r2 + a2;
    const F1 denomH = ffma(r2, r2, ya * y);
    const F1 inv_s = frcp(s);
    const F1 inv_r = frcp(r);
    const F1 inv_denom = frcp(denom);
    const F1 inv_denomH = frcp(denomH);
    const F1 Nx = ffma(r, x, -(a * z));
    const F1 Nz = ffma(r, z, a * x);
    const F1 lx = Nx * inv_denom, ly = y * inv_r, lz = Nz * inv_denom;
    const F1 r3 = r2 * r;
    const F1 H = (M * r3) * inv_denomH;
    const F1 Q = ffma(lx, px, lz * pz);
    const F1 lp = ffma(ly, py, Q) - pt;
    const F1 c = (2.0f * H) * lp;
    const F1 fx = px - c * lx, fy = py - c * ly, fz = pz - c * lz;
    const F1 inv_2r = 0.5f * inv_r;
    const F1 w1 = (u * inv_s + 1.0f) * inv_2r;
    const F1 w2 = ((u + 2.0f * a2) * inv_s + 1.0f) * inv_2r;
    const F1 XP = ffma(x, px, z * pz);
    const F1 inv_r2 = inv_r * inv_r;
    const F1 G = ffma(-2.0f * r, Q, XP) * inv_denom - (y * py) * inv_r2;
    const F1 inv_denomH2 = inv_denomH * inv_denomH;
    const F1 HR = ((M * r2) * inv_denomH2) * (3.0f * (ya * y) - r2 * r2);
    const F1 HY = -(2.0f * M) * ya * (r3 * inv_denomH2);
    const F1 lp2 = lp * lp;
    const F1 Z = ffma(HR, lp2, c * G);
    const F1 cD = c * inv_denom;
    const F1 fpx = ffma(x * w1, Z, cD * ffma(r, px, a * pz));
    const F1 fpy = ffma(y * w2, Z, ffma(HY, lp2, (c * py) * inv_r));
    const F1 fpz = ffma(z * w1, Z, cD * ffma(r, pz, -(a * px)));
    lamY.x -= fpx.d;
    lamY.y -= fpy.d;
    lamY.z -= fpz.d;
    lamY.px += fx.d;
    lamY.py += fy.d;
    lamY.pz += fz.d;
    lam_pt += 2.0f * fmaf(H.d, lp.v, H.v * lp.d);
    const F1 r2_a = a * ((2.0f * yy - u) * inv_s - 1.0f);
    const F1 r_a = r2_a * inv_2r;
    const F1 D_a = r2_a + 2.0f * a;
    const F1 inv_denom2 = inv_denom * inv_denom;
    const F1 lx_a = ((r_a * x - z) * denom - Nx * D_a) * inv_denom2;
    const F1 lz_a = ((r_a * z + x) * denom - Nz * D_a) * inv_denom2;
    const F1 ly_a = -(y * r_a) * inv_r2;
    const F1 DH_a = 2.0f * r2 * r2_a + (2.0f * a) * yy;
    const F1 H_a = (M * inv_denomH2) * ((3.0f * r2) * denomH * r_a - r3 * DH_a);
    const F1 lp_a = ffma(lx_a, px, ffma(ly_a, py, lz_a * pz));
    const F1 ham_a = -(H_a * lp2 + 2.0f * H * lp * lp_a);
    lam_a += ham_a.d;
}

// The exact right-hand side in FP32 (the primal of rhs_vjp32).  Outside
// the horizon r > |a| and u = rho^2 - a^2 >= r^2 - a^2 > 0, so the radius
// solve u + sqrt(u^2 + 4 a^2 y^2) does not cancel in FP32.
KT_HDI V6f rhs_val32(const V6f& Y, float pt, float a, float a2, float M) {
    const float x = Y.x, y = Y.y, z = Y.z, px = Y.px, py = Y.py, pz = Y.pz;
    const float rho2 = fmaf(x, x, fmaf(y, y, z * z));
    const float u = rho2 - a2;
    const float ya = a2 * y;
    const float s_sq = fmaf(u, u, 4.0f * (ya * y));
    const float s = s_sq * rsq32(s_sq);
    const float r2 = 0.5f * (u + s);
    const float r = r2 * rsq32(r2);
    const float denom = r2 + a2;
    const float denomH = fmaf(r2, r2, ya * y);
    const float inv_s = rcp32(s), inv_r = rcp32(r), inv_denom = rcp32(denom), inv_denomH = rcp32(denomH);
    const float Nx = fmaf(r, x, -(a * z)), Nz = fmaf(r, z, a * x);
    const float lx = Nx * inv_denom, ly = y * inv_r, lz = Nz * inv_denom;
    const float r3 = r2 * r;
    const float H = (M * r3) * inv_denomH;
    const float Q = fmaf(lx, px, lz * pz);
    const float lp = fmaf(ly, py, Q) - pt;
    const float c = (2.0f * H) * lp;
    const float inv_2r = 0.5f * inv_r;
    const float w1 = fmaf(u, inv_s, 1.0f) * inv_2r;
    const float w2 = fmaf(u + 2.0f * a2, inv_s, 1.0f) * inv_2r;
    const float XP = fmaf(x, px, z * pz);
    const float inv_r2 = inv_r * inv_r;
    const float G = fmaf(fmaf(-2.0f * r, Q, XP), inv_denom, -((y * py) * inv_r2));
    const float inv_denomH2 = inv_denomH * inv_denomH;
    const float HR = ((M * r2) * inv_denomH2) * (3.0f * (ya * y) - r2 * r2);
    const float HY = -(2.0f * M) * ya * (r3 * inv_denomH2);
    const float lp2 = lp * lp;
    const float Zz = fmaf(HR, lp2, c * G);
    const float cD = c * inv_denom;
    V6f d;
    d.x = fmaf(-c, lx, px);
    d.y = fmaf(-c, ly, py);
    d.z = fmaf(-c, lz, pz);
    d.px = fmaf(x * w1, Zz, cD * fmaf(r, px, a * pz));
    d.py = fmaf(y * w2, Zz, fmaf(HY, lp2, (c * py) * inv_r));
    d.pz = fmaf(z * w1, Zz, cD * fmaf(r, pz, -(a * px)));
    return d;
}

// The inputs of stages 2..6 of a step from s, in FP32: the linearisation
// points of the FP32 step reversal, within a few FP32 ulps of the forward
// pass's FP64 stage inputs (the reversal's own FP32 rounding is larger).
KT_HDI void dp_stages32(const V6f& s, float hs, float pt, float a, float a2, float M, V6f* Y) {
#define KTA_F(M_) V6f{M_(x), M_(y), M_(z), M_(px), M_(py), M_(pz)}
    const V6f k1 = rhs_val32(s, pt, a, a2, M);
#define KTA_F2(c) fmaf(hs, (float)KTA_A21 * k1.c, s.c)
    Y[0] = KTA_F(KTA_F2);
    const V6f k2 = rhs_val32(Y[0], pt, a, a2, M);
#define KTA_F3(c) fmaf(hs, fmaf((float)KTA_A32, k2.c, (float)KTA_A31 * k1.c), s.c)
    Y[1] = KTA_F(KTA_F3);
    const V6f k3 = rhs_val32(Y[1], pt, a, a2, M);
#define KTA_F4(c) fmaf(hs, fmaf((float)KTA_A43, k3.c, fmaf((float)KTA_A42, k2.c, (float)KTA_A41 * k1.c)), s.c)
    Y[2] = KTA_F(KTA_F4);
    const V6f k4 = rhs_val32(Y[2], pt, a, a2, M);
#define KTA_F5(c) \
    fmaf(hs, fmaf((float)KTA_A54, k4.c, fmaf((float)KTA_A53, k3.c, fmaf((float)KTA_A52, k2.c, (float)KTA_A51 * k1.c))), s.c)
    Y[3] = KTA_F(KTA_F5);
    const V6f k5 = rhs_val32(Y[3], pt, a, a2, M);
#define KTA_F6(c)                                                                                                  \
    fmaf(hs,                                                                                                       \
         fmaf((float)KTA_A65, k5.c,                                                                                \
              fmaf((float)KTA_A64, k4.c, fmaf((float)KTA_A63, k3.c, fmaf((float)KTA_A62, k2.c, (float)KTA_A61 * k1.c)))), \
         s.c)
    Y[4] = KTA_F(KTA_F6);
#undef KTA_F
#undef KTA_F2
#undef KTA_F3
#undef KTA_F4
#undef KTA_F5
#undef KTA_F6
}

// dp_step_reverse in FP32 (same recursion; see dp_step_reverse).
KT_HDI V6f dp_step_reverse32(const V6f& s, const V6f& ns, const V6f* Y, float hs, float pt, float a, float a2,
                             float M, V6f lam_ns, const V6f* lk, const V6f& lam_s_dir, float& lam_pt, float& lam_a) {
    if (lk) rhs_vjp32(ns, pt, a, a2, M, lk[5], lam_ns, lam_pt, lam_a);  // k7 = f(ns)
    V6f l1 = v6f_zero(), l2 = v6f_zero(), l3 = v6f_zero(), l4 = v6f_zero(), l5 = v6f_zero(), l6 = v6f_zero();
    v6f_axpy(l1, hs * (float)KTA_B1, lam_ns);
    v6f_axpy(l3, hs * (float)KTA_B3, lam_ns);
    v6f_axpy(l4, hs * (float)KTA_B4, lam_ns);
    v6f_axpy(l5, hs * (float)KTA_B5, lam_ns);
    v6f_axpy(l6, hs * (float)KTA_B6, lam_ns);
    if (lk) {
        v6f_add(l1, lk[0]);
        v6f_add(l3, lk[1]);
        v6f_add(l4, lk[2]);
        v6f_add(l5, lk[3]);
        v6f_add(l6, lk[4]);
    }
    V6f ls = lam_ns;
    v6f_add(ls, lam_s_dir);
    V6f lY;
    lY = v6f_zero();
    rhs_vjp32(Y[4], pt, a, a2, M, l6, lY, lam_pt, lam_a);
    v6f_add(ls, lY);
    v6f_axpy(l1, hs * (float)KTA_A61, lY);
    v6f_axpy(l2, hs * (float)KTA_A62, lY);
    v6f_axpy(l3, hs * (float)KTA_A63, lY);
    v6f_axpy(l4, hs * (float)KTA_A64, lY);
    v6f_axpy(l5, hs * (float)KTA_A65, lY);
    lY = v6f_zero();
    rhs_vjp32(Y[3], pt, a, a2, M, l5, lY, lam_pt, lam_a);
    v6f_add(ls, lY);
    v6f_axpy(l1, hs * (float)KTA_A51, lY);
    v6f_axpy(l2, hs * (float)KTA_A52, lY);
    v6f_axpy(l3, hs * (float)KTA_A53, lY);
    v6f_axpy(l4, hs * (float)KTA_A54, lY);
    lY = v6f_zero();
    rhs_vjp32(Y[2], pt, a, a2, M, l4, lY, lam_pt, lam_a);
    v6f_add(ls, lY);
    v6f_axpy(l1, hs * (float)KTA_A41, lY);
    v6f_axpy(l2, hs * (float)KTA_A42, lY);
    v6f_axpy(l3, hs * (float)KTA_A43, lY);
    lY = v6f_zero();
    rhs_vjp32(Y[1], pt, a, a2, M, l3, lY, lam_pt, lam_a);
    v6f_add(ls, lY);
    v6f_axpy(l1, hs * (float)KTA_A31, lY);
    v6f_axpy(l2, hs * (float)KTA_A32, lY);
    lY = v6f_zero();
    rhs_vjp32(Y[0], pt, a, a2, M, l2, lY, lam_pt, lam_a);
    v6f_add(ls, lY);
    v6f_axpy(l1, hs * (float)KTA_A21, lY);
    rhs_vjp32(s, pt, a, a2, M, l1, ls, lam_pt, lam_a);
    return ls;
}

// ------------------------------------------------- crossing (dense output)
// Hit parameter t and interpolated hit state (x, z, px, py, pz) of a step
// that crossed y = 0, as the tracer computes them.
struct Crossing {
    double t;
    double x, z, px, py, pz;
};
KT_HDI Crossing crossing_hit(const V6& s, const V6& ns, const V6& k1, const V6& k7, const V6& dsum, double hs,
                             int disk_root_mode) {
    Crossing c;
    double t;
    if (disk_root_mode == 1) {
        t = disk_root_s(s.y, ns.y, k1.y, k7.y, dsum.y, hs);
    } else {
        double den = s.y - ns.y;
        t = den != 0.0 ? s.y / den : 0.5;
        t = fmin(fmax(t, 0.0), 1.0);
    }
    c.t = t;
    c.x = dense_s(s.x, ns.x, k1.x, k7.x, dsum.x, hs, t);
    c.z = dense_s(s.z, ns.z, k1.z, k7.z, dsum.z, hs, t);
    c.px = dense_s(s.px, ns.px, k1.px, k7.px, dsum.px, hs, t);
    c.py = dense_s(s.py, ns.py, k1.py, k7.py, dsum.py, hs, t);
    c.pz = dense_s(s.pz, ns.pz, k1.pz, k7.pz, dsum.pz, hs, t);
    return c;
}

// Weights of the DOPRI5 interpolant at t on its inputs:
//   P(t) = w0 y0 + w1 y1 + h (wk1 k1 + sum_{i=3..6} d_i w5 k_i + wk7 k7).
struct DenseW {
    double w0, w1, wk1, wk7, w5;
};
KT_HDI DenseW dense_weights(double t) {
    double p2 = t, p3 = t - t * t, p4 = t * t - t * t * t, p5 = t * t - 2.0 * t * t * t + t * t * t * t;
    DenseW w;
    w.w0 = 1.0 - p2 + p3 - 2.0 * p4;
    w.w1 = p2 - p3 + 2.0 * p4;
    w.wk1 = p3 - p4 + Dense::d1 * p5;
    w.wk7 = -p4 + Dense::d7 * p5;
    w.w5 = p5;
    return w;
}

// Adjoint of the crossing: lam_hit (x, z, px, py, pz) -> direct adjoints of
// s (lam_s), ns (lam_ns) and k1, k3, k4, k5, k6, k7 (lk[0..5]), including
// the root's implicit dependence on the y-components.
KT_HDI void crossing_reverse(const V6& s, const V6& ns, const V6& k1, const V6& k7, const V6& dsum, double hs,
                             int disk_root_mode, double t, const double lam_hit[5], V6& lam_s, V6& lam_ns, V6* lk) {
    const DenseW w = dense_weights(t);
    // d hit / d t from the interpolant's derivative
    double dxdt = dense_s_dt(s.x, ns.x, k1.x, k7.x, dsum.x, hs, t);
    double dzdt = dense_s_dt(s.z, ns.z, k1.z, k7.z, dsum.z, hs, t);
    double dpxdt = dense_s_dt(s.px, ns.px, k1.px, k7.px, dsum.px, hs, t);
    double dpydt = dense_s_dt(s.py, ns.py, k1.py, k7.py, dsum.py, hs, t);
    double dpzdt = dense_s_dt(s.pz, ns.pz, k1.pz, k7.pz, dsum.pz, hs, t);
    double lam_t = lam_hit[0] * dxdt + lam_hit[1] * dzdt + lam_hit[2] * dpxdt + lam_hit[3] * dpydt +
                   lam_hit[4] * dpzdt;
    const double di[4] = {Dense::d3, Dense::d4, Dense::d5, Dense::d6};
    // components carried by the hit state
#define KTA_CR(C, L)                              \
    lam_s.C += (L) * w.w0;                        \
    lam_ns.C += (L) * w.w1;                       \
    lk[0].C += (L) * hs * w.wk1;                  \
    lk[5].C += (L) * hs * w.wk7;                  \
    for (int i = 0; i < 4; ++i) lk[1 + i].C += (L) * hs * di[i] * w.w5;
    KTA_CR(x, lam_hit[0])
    KTA_CR(z, lam_hit[1])
    KTA_CR(px, lam_hit[2])
    KTA_CR(py, lam_hit[3])
    KTA_CR(pz, lam_hit[4])
    // the root: y(t; inputs) = 0 (true root) or the secant t = y0 / (y0 - y1)
    if (disk_root_mode == 1) {
        double dydt = dense_s_dt(s.y, ns.y, k1.y, k7.y, dsum.y, hs, t);
        if (dydt != 0.0) {
            double q = -lam_t / dydt;
            KTA_CR(y, q)
        }
    } else {
        double den = s.y - ns.y;
        if (den != 0.0) {
            double tr = s.y / den;
            if (tr > 0.0 && tr < 1.0) {
                double inv2 = 1.0 / (den * den);
                lam_s.y += lam_t * (-ns.y * inv2);
                lam_ns.y += lam_t * (s.y * inv2);
            }
        }
    }
#undef KTA_CR
}

// --------------------------------------------------- shading (dual pass)
// Per-render constants of the differentiable scene.
struct GradConsts {
    const double* dlut;    // d disk_lut[i] / d spin
    double d_rin_da;       // d disk_inner / d spin
    double dTs_da;         // d temperature_scale / d spin
    double dTs_dm;         // d temperature_scale / d mdot
    double incl, r_obs;    // theta[1], theta[2]
    double fov_scale;
    double w_scale;        // exposure / spp: d hdr / d colour
};

// Emission of a crossing in one forward-mode direction: c = d_col * alpha
// and alpha, with the tracer's gates (disk range, disc_denom > 0, finite
// positive temperature, LUT range).  a, aa, rin, Tsc carry the direction's
// tangents of spin, spin^2, disk_inner, temperature_scale; lut_tan scales
// the d LUT / d spin table (1 in the spin direction, else 0).
struct ShadeD {
    D1 c[3];
    D1 alpha;
    bool lit;
};

KT_HDI D1 gradient_noise_d(D1 u, D1 v, uint64_t seed_mix) {
    double fu = floor(u.v), fv = floor(v.v);
    int ix0 = (int)fu, iy0 = (int)fv;
    D1 fx = u - (double)ix0;
    D1 fy = v - (double)iy0;
    double g00x, g00y, g10x, g10y, g01x, g01y, g11x, g11y;
    lattice_gradient(ix0, iy0, seed_mix, g00x, g00y);
    lattice_gradient(ix0 + 1, iy0, seed_mix, g10x, g10y);
    lattice_gradient(ix0, iy0 + 1, seed_mix, g01x, g01y);
    lattice_gradient(ix0 + 1, iy0 + 1, seed_mix, g11x, g11y);
    D1 n00 = g00x * fx + g00y * fy;
    D1 n10 = g10x * (fx - 1.0) + g10y * fy;
    D1 n01 = g01x * fx + g01y * (fy - 1.0);
    D1 n11 = g11x * (fx - 1.0) + g11y * (fy - 1.0);
    D1 u_f = fx * fx * fx * (fx * (fx * 6.0 - 15.0) + 10.0);
    D1 v_f = fy * fy * fy * (fy * (fy * 6.0 - 15.0) + 10.0);
    D1 x1 = dfma(u_f, n10 - n00, n00);
    D1 x2 = dfma(u_f, n11 - n01, n01);
    return dfma(v_f, x2 - x1, x1);
}
KT_HDI D1 fbm_d(D1 u, D1 v, int octaves, uint64_t seed_mix) {
    if (octaves <= 0) return cst(0.5);
    D1 sum = cst(0.0);
    double amp = 1.0, freq = 1.0, amp_sum = 0.0;
    for (int i = 0; i < octaves; ++i) {
        sum = sum + gradient_noise_d(u * freq, v * freq, seed_mix) * amp;
        amp_sum += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    D1 vv = amp_sum > 0.0 ? sum / amp_sum : cst(0.0);
    return vv * 0.5 + 0.5;
}

KT_HDI ShadeD shade_d(const RenderParams& P, const GradConsts& G, D1 x_hit, D1 z_hit, D1 px, D1 py, D1 pz, D1 pt,
                      D1 kt, D1 a, D1 aa, D1 rin, D1 Tsc, double lut_tan) {
    ShadeD o;
    o.lit = false;
    const double M = P.M;
    // Kerr-Schild radius at y = 0
    D1 rho2 = dfma(x_hit, x_hit, z_hit * z_hit);
    D1 u0 = rho2 - aa;
    D1 r_hit = dsqrt(0.5 * (u0 + dsqrt(u0 * u0)));
    if (!(r_hit.v >= rin.v && r_hit.v <= P.disk_outer)) return o;
    const double sqrt_M = sqrt(M);
    D1 r_sqrt = dsqrt(r_hit);
    D1 disc_denom = 1.0 - 3.0 * M / r_hit + (2.0 * a) * sqrt_M / (r_hit * r_sqrt);
    if (!(disc_denom.v > 0.0)) return o;
    D1 omega = -sqrt_M / (r_hit * r_sqrt + a * sqrt_M);
    D1 r2_hit = r_hit * r_hit;
    D1 inv_dh = 1.0 / (r2_hit + aa);
    D1 hlx = (r_hit * x_hit - a * z_hit) * inv_dh;
    D1 hlz = (r_hit * z_hit + a * x_hit) * inv_dh;
    D1 ldotp = hlx * px + hlz * pz;
    D1 ldotxi = hlx * z_hit - hlz * x_hit;
    D1 Hh = M * r_hit * r2_hit * (1.0 / (r2_hit * r2_hit));
    D1 L = (z_hit * px - x_hit * pz) + 2.0 * Hh * ldotxi * (ldotp + kt);
    D1 l_over_e = L / (-pt);
    D1 g = 1.0 / ((1.0 / dsqrt(disc_denom)) * (1.0 - omega * l_over_e));
    D1 uu = (r_hit - rin) / (P.disk_outer - rin);
    // disk LUT at texel u*N - 0.5, clamped
    D1 xb = uu * (double)P.lut_n - 0.5;
    double fl = floor(xb.v);
    int i0 = (int)fl;
    D1 wl = xb - fl;
    int i1 = i0 + 1;
    i0 = i0 < 0 ? 0 : (i0 > P.lut_n - 1 ? P.lut_n - 1 : i0);
    i1 = i1 < 0 ? 0 : (i1 > P.lut_n - 1 ? P.lut_n - 1 : i1);
    D1 c0{P.disk_lut[i0], lut_tan * G.dlut[i0]}, c1{P.disk_lut[i1], lut_tan * G.dlut[i1]};
    D1 sampled = c0 + wl * (c1 - c0);
    D1 T = Tsc * sampled * g;
    if (P.noise_detail > 0) {
        // Boyer-Lindquist phi at y = 0 and the radius r2 / r as the tracer forms it
        D1 s_abs = dsqrt(u0 * u0);
        D1 r2 = 0.5 * (u0 + s_abs);
        D1 r_bl = r2 / dsqrt(r2);
        D1 inv = 1.0 / (r2 + aa);
        D1 cos_phi = (r_bl * z_hit + a * x_hit) * inv;
        D1 sin_phi = (r_bl * x_hit - a * z_hit) * inv;
        D1 phi = datan2(sin_phi, cos_phi);
        D1 om = 1.0 / (a + dsqrt(r_bl * r_bl * r_bl * P.inv_M));
        D1 phase = phi + P.noise_winding * om;
        D1 sn, cs;
        dsincos(phase, sn, cs);
        D1 nu = r_bl * cs * P.noise_scale;
        D1 nv = r_bl * sn * P.noise_scale;
        D1 nval = fbm_d(nu, nv, P.noise_detail, P.noise_seed_mix);
        T = T * (1.0 + P.noise_strength * (nval - 0.5));
    }
    if (!(__builtin_isfinite(T.v) && T.v > 0.0)) return o;
    if (T.v > P.lut_max_temp) return o;  // overflow: black, the frame is rejected
    D1 pos = T / P.lut_max_temp * (double)(P.lut_n - 1);
    double flb = floor(pos.v);
    int j0 = (int)flb;
    D1 wb = pos - flb;
    int j1 = j0 + 1;
    j0 = j0 < 0 ? 0 : (j0 > P.lut_n - 1 ? P.lut_n - 1 : j0);
    j1 = j1 < 0 ? 0 : (j1 > P.lut_n - 1 ? P.lut_n - 1 : j1);
    D1 rgb[3];
    for (int k = 0; k < 3; ++k) {
        double b0 = P.bb_lut[j0 * 3 + k], b1 = P.bb_lut[j1 * 3 + k];
        rgb[k] = b0 + wb * (b1 - b0);
    }
    D1 luma = 0.2126 * rgb[0] + 0.7152 * rgb[1] + 0.0722 * rgb[2];
    D1 alpha = 1.0 - dexp(-(luma / P.wl_step) * 0.5);
    for (int k = 0; k < 3; ++k) o.c[k] = rgb[k] * alpha;
    o.alpha = alpha;
    o.lit = true;
    return o;
}

// Directions of the shading pass: the hit state (5), pt, kt, spin, mdot.
// Returns the adjoints of the hit state in lam_hit and accumulates those of
// pt, kt, spin, mdot, given lam_c (of c) and lam_alpha.
KT_HDI void shade_reverse(const RenderParams& P, const GradConsts& G, const Crossing& h, double pt, double kt,
                          const double lam_c[3], double lam_alpha, double lam_hit[5], double& lam_pt, double& lam_kt,
                          double& g_a, double& g_m) {
    const double vals[7] = {h.x, h.z, h.px, h.py, h.pz, pt, kt};
    for (int dir = 0; dir < 9; ++dir) {
        D1 in[7];
        for (int k = 0; k < 7; ++k) in[k] = D1{vals[k], dir == k ? 1.0 : 0.0};
        const bool da = dir == 7, dm = dir == 8;
        D1 a{P.a, da ? 1.0 : 0.0};
        D1 aa{P.aa, da ? 2.0 * P.a : 0.0};
        D1 rin{P.disk_inner, da ? G.d_rin_da : 0.0};
        D1 Tsc{P.T_scale, da ? G.dTs_da : (dm ? G.dTs_dm : 0.0)};
        ShadeD o = shade_d(P, G, in[0], in[1], in[2], in[3], in[4], in[5], in[6], a, aa, rin, Tsc, da ? 1.0 : 0.0);
        double d = 0.0;
        if (o.lit) d = lam_c[0] * o.c[0].d + lam_c[1] * o.c[1].d + lam_c[2] * o.c[2].d + lam_alpha * o.alpha.d;
        if (dir < 5) lam_hit[dir] = d;
        else if (dir == 5) lam_pt += d;
        else if (dir == 6) lam_kt += d;
        else if (da) g_a += d;
        else g_m += d;
    }
}

// Values of a crossing's emission (the tracer's shade_crossing in the form
// the reverse pass needs: c = d_col * alpha and alpha).
KT_HDI ShadeD shade_value(const RenderParams& P, const GradConsts& G, const Crossing& h, double pt, double kt) {
    return shade_d(P, G, cst(h.x), cst(h.z), cst(h.px), cst(h.py), cst(h.pz), cst(pt), cst(kt), cst(P.a),
                   cst(P.aa), cst(P.disk_inner), cst(P.T_scale), 0.0);
}

// -------------------------------------------------- ray start (dual pass)
// Initial state, pt and kt of pixel (px, py) (sample 0) in one direction of
// theta: dir 0 spin, 1 inclination, 2 observer distance.  The camera is the
// observer's: pos = r (sin i, cos i, 0), fwd = -pos / r, rgt = (0, 0, -1),
// up = (-cos i, sin i, 0).
struct StartD {
    D1 s[6];
    D1 pt, kt;
};
KT_HDI StartD start_d(const RenderParams& P, const GradConsts& G, int pix_x, int pix_y, int dir) {
    D1 a{P.a, dir == 0 ? 1.0 : 0.0};
    D1 aa = a * a;
    D1 incl{G.incl, dir == 1 ? 1.0 : 0.0};
    D1 r_obs{G.r_obs, dir == 2 ? 1.0 : 0.0};
    D1 si, ci;
    dsincos(incl, si, ci);
    uint32_t pixel_index = (uint32_t)(pix_y * P.W + pix_x);
    uint64_t base = (uint64_t)pixel_index << 32;
    double rx = (double)pcg32(base) * 0x1p-32;
    double ry = (double)pcg32(base ^ 0xda3e39cb94b95bdbULL) * 0x1p-32;
    double u = (((double)pix_x + rx) * P.inv_w_2 - 1.0) * P.aspect * G.fov_scale;
    double v = (1.0 - ((double)pix_y + ry) * P.inv_h_2) * G.fov_scale;
    // dir = u rgt + v up + fwd
    D1 dx = v * (-ci) - si;
    D1 dy = v * si - ci;
    D1 dz = cst(-u);
    D1 inv_n = 1.0 / dsqrt(dx * dx + dy * dy + dz * dz);
    D1 nx = dx * inv_n, ny = dy * inv_n, nz = dz * inv_n;
    D1 x = r_obs * si, y = r_obs * ci, z = cst(0.0);
    D1 rho2 = x * x + y * y;
    D1 uu = rho2 - aa;
    D1 r2 = 0.5 * (uu + dsqrt(uu * uu + 4.0 * aa * y * y));
    D1 r = dsqrt(r2);
    D1 inv_denom = 1.0 / (r2 + aa);
    D1 lx = (r * x - a * z) * inv_denom;
    D1 ly = y / r;
    D1 lz = (r * z + a * x) * inv_denom;
    D1 Hs = P.M * r2 * r / (r2 * r2 + aa * y * y);
    D1 ldotn = lx * nx + ly * ny + lz * nz;
    D1 g_tt = -1.0 + 2.0 * Hs;
    D1 g_tn = 2.0 * Hs * ldotn;
    D1 g_nn = 1.0 + 2.0 * Hs * ldotn * ldotn;
    D1 disc = g_tn * g_tn - g_tt * g_nn;
    StartD o;
    o.kt = (-g_tn - dsqrt(disc)) / g_tt;
    o.pt = g_tt * o.kt + g_tn;
    D1 sc = o.kt + ldotn;
    o.s[0] = x;
    o.s[1] = y;
    o.s[2] = z;
    o.s[3] = nx + 2.0 * Hs * lx * sc;
    o.s[4] = ny + 2.0 * Hs * ly * sc;
    o.s[5] = nz + 2.0 * Hs * lz * sc;
    return o;
}

}  // namespace kt
