// Device math of the FP64 tracer (gfx950).
//
// Every trajectory operation is an IEEE-exact binary64 +, -, *, /, sqrt or
// fma, in a fixed order (the library is built with -ffp-contract=off, so
// a*b+c is never fused behind the source's back); the only non-elementary
// trajectory function is pow_m02, written below from its mathematics.  The
// tracer therefore reproduces any binary64 evaluation of the same operation
// sequence bit for bit.  The noise's lattice gradients use turn_cos_sin
// below, written from its mathematics.  The other shading-only
// transcendentals (exp, sin, cos, atan2, asin) come from the device math
// library and differ from a host libm by at most a few ulp; they never feed
// back into a trajectory.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kt_internal.hpp"

// FP64 routines below are host+device so the adjoint's building blocks can
// be exercised by a host build of the same source.
#define KT_HDI __host__ __device__ __forceinline__

namespace kt {

// cuda/kerr_math.cuh:26-32: one 64-bit LCG step then the XSH-RR output.
KT_HDI uint32_t pcg32(uint64_t state) {
    uint64_t x = state * 6364136223846793005ULL + 1442695040888963407ULL;
    uint32_t xs = (uint32_t)(((x >> 18u) ^ x) >> 27u);
    uint32_t rot = (uint32_t)(x >> 59u);
    return (xs >> rot) | (xs << ((32u - rot) & 31u));
}

// x^(-1/5) for the step-size controller (kerr_trace.cuh:119 uses __powf).
// x = m 2^e, m in [1,2), e = 5q + j: x^(-1/5) = m^(-1/5) 2^(-j/5) 2^(-q).
// m^(-1/5): quadratic start (|rel. error| < 2e-3), then four Newton steps
// y <- y (6 - m y^5) / 5 on y^-5 = m (error 2e-3 -> 1e-5 -> 4e-10 -> 5e-19).
// Branch-free: the main path runs for every x and the special values
// (NaN, 0, infinities, negatives) are selected at the end, so a wave takes
// no branches here; q = floor(e / 5) comes from an unsigned division of
// e + 1075 >= 1 (e >= -1074 for subnormals).
KT_HDI double pow_m02(double x) {
    int e = 0;
    double m = frexp(x, &e);
    m = m * 2.0;
    e = e - 1;
    const int q = (int)((unsigned)(e + 1075) / 5u) - 215;
    const int j = e - 5 * q;
    double tj = 1.0;
    tj = j == 1 ? 0x1.bdb8cdadbe120p-1 : tj;   // 2^(-1/5)
    tj = j == 2 ? 0x1.8406003b2ae5cp-1 : tj;   // 2^(-2/5)
    tj = j == 3 ? 0x1.51cb453b9536cp-1 : tj;   // 2^(-3/5)
    tj = j == 4 ? 0x1.2611186bae675p-1 : tj;   // 2^(-4/5)
    double y = fma(fma(0.0517578125, m, -0.283203125), m, 1.2294921875);
#pragma unroll
    for (int it = 0; it < 4; ++it) {
        double y2 = y * y;
        double y5 = y2 * y2 * y;
        y = y * (6.0 - m * y5) * 0.2;
    }
    double out = ldexp(y * tj, -q);
    out = x < 0.0 ? __builtin_nan("") : out;
    out = x == __builtin_huge_val() ? 0.0 : out;
    out = x == 0.0 ? __builtin_huge_val() : out;
    out = x != x ? x : out;
    return out;
}

// Correctly rounded FP64 square root and quotient for the operands the
// tracer produces: positive, normal, exponents far from the extremes
// (radii, squared radii, products of radii and step errors).  On the device
// they are the refinement cores of the full library sequences without the
// operand scaling and special-value fix-ups those wrap around them:
//   sqrt: y = rsq(x), g = x y, h = y / 2, r = 1/2 - g h, g += g r,
//         h += h r, then twice d = x - g^2, g += d h  (Goldschmidt);
//   n / d: r = rcp(d), twice e = 1 - d r, r += r e; q = n r,
//          q += (n - d q) r  (Markstein).
// For such operands the library sequences run exactly these operations,
// so the results are the same correctly rounded values; the host build
// uses the library functions.
KT_HDI double sqrt_pos(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    const double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    g = fma(d, h, g);
    return g;
#else
    return sqrt(x);
#endif
}
KT_HDI double div_pos(double n, double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(d);
    double e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    e = fma(-d, r, 1.0);
    r = fma(r, e, r);
    const double q = n * r;
    const double rem = fma(-d, q, n);
    return fma(rem, r, q);
#else
    return n / d;
#endif
}

// cuda/kerr_math.cuh:50-59: Kerr-Schild radius from r^4 - (rho^2 - a^2) r^2 - a^2 y^2 = 0
// (s_sq * rsqrt(s_sq) evaluated as one square root).
KT_HDI double ks_r(double x, double y, double z, double a2) {
    double rho2 = fma(x, x, fma(y, y, z * z));
    double u = rho2 - a2;
    double s_sq = fma(u, u, 4.0 * (a2 * y) * y);
    double r2 = 0.5 * (u + sqrt_pos(s_sq));
    return sqrt_pos(r2);
}

struct V6 {
    double x, y, z, px, py, pz;
};

// Geometry of one render that the right-hand side reads.
struct Geo {
    double a, aa, M;
    bool exact;  // exact Hamiltonian dr^2/dy instead of the as-written one
    // dr/dy weight w2 = fma(u + w2c, w2f / s, 1) / (2 r): exact form
    // (w2c, w2f) = (2 a^2, 1), the reference's as-written form (4 a^2, 1/2);
    // both products are exact scalings, so either form rounds as written
    double w2c, w2f;
};
KT_HDI Geo make_geo(double a, double aa, double M, bool exact) {
    return Geo{a, aa, M, exact, exact ? 2.0 * aa : 4.0 * aa, exact ? 1.0 : 0.5};
}

// cuda/kerr_math.cuh:82-156: Hamiltonian right-hand side in Cartesian
// Kerr-Schild coordinates, dx/dl = p - 2H (l.p - pt) l and
// dp/dl = grad(H) (l.p - pt)^2 + 2H (l.p - pt) grad(l.p), with the position
// derivatives in factored form: dr/dx_j = (x w1, y w2, z w1),
// grad(l.p) = dr/dx_j G + (r px + a pz, py / r, r pz - a px) terms and
// grad(H) = dr/dx_j HR + (0, HY, 0) (the reference's as-written dr/dy,
// whose u/s term is halved, enters through w2).  Products that feed a sum
// are fused with explicit fma; the reciprocals 1/s, 1/r, 1/(r^2+a^2),
// 1/(r^4+a^2y^2) come from one division.  r_out (optional) receives the
// Kerr-Schild radius of st, the same value ks_r() returns.  a^2y y is formed
// once, for 4 a^2y^2 and for r^4 + a^2y^2, and 2 H as (2M r^3)/(r^4 + a^2y^2)
// (scalings by 2 and 4 are exact, so these round as (4 a^2y) y and
// 2 (M r^3/(r^4 + a^2y^2)) would).
KT_HDI V6 rhs(const V6& st, double pt, const Geo& g, double* r_out = nullptr) {
    const double x = st.x, y = st.y, z = st.z, px = st.px, py = st.py, pz = st.pz;
    const double a2 = g.aa, a = g.a;
    double rho2 = fma(x, x, fma(y, y, z * z));
    double u = rho2 - a2;
    double ya = a2 * y;
    double yay = ya * y;
    double s_sq = fma(u, u, 4.0 * yay);
    double s = sqrt_pos(s_sq);
    double r2 = 0.5 * (u + s);
    double r = sqrt_pos(r2);
    if (r_out) *r_out = r;
    double denom = r2 + a2;
    double denomH = fma(r2, r2, yay);
    double sr = s * r;
    double dd = denom * denomH;
    double q = div_pos(1.0, sr * dd);
    double inv_s = (r * dd) * q;
    double inv_r = (s * dd) * q;
    double inv_denom = (sr * denomH) * q;
    double inv_denomH = (sr * denom) * q;
    double Nx = fma(r, x, -(a * z));
    double Nz = fma(r, z, a * x);
    double lx = Nx * inv_denom;
    double ly = y * inv_r;
    double lz = Nz * inv_denom;
    double r3 = r2 * r;
    double H2 = ((2.0 * g.M) * r3) * inv_denomH;  // 2 H
    double lp = fma(lx, px, fma(ly, py, fma(lz, pz, -pt)));
    double c = H2 * lp;
    V6 d;
    d.x = fma(-c, lx, px);
    d.y = fma(-c, ly, py);
    d.z = fma(-c, lz, pz);
    double inv_2r = 0.5 * inv_r;
    double w1 = fma(u, inv_s, 1.0) * inv_2r;
    double w2 = fma(u + g.w2c, g.w2f * inv_s, 1.0) * inv_2r;
    double Q = fma(lx, px, lz * pz);
    double XP = fma(x, px, z * pz);
    double inv_r2 = inv_r * inv_r;
    double G = fma(fma(-2.0 * r, Q, XP), inv_denom, -((y * py) * inv_r2));
    double inv_denomH2 = inv_denomH * inv_denomH;
    double HR = ((g.M * r2) * inv_denomH2) * fma(3.0 * ya, y, -(r2 * r2));
    double HY = -((2.0 * g.M) * ya) * (r3 * inv_denomH2);
    double lp2 = lp * lp;
    double Z = fma(HR, lp2, c * G);
    double cD = c * inv_denom;
    d.px = fma(x * w1, Z, cD * fma(r, px, a * pz));
    d.py = fma(y * w2, Z, fma(HY, lp2, (c * py) * inv_r));
    d.pz = fma(z * w1, Z, cD * fma(r, pz, -(a * px)));
    return d;
}

// Mixed-precision right-hand side: FP64 state in, FP64 derivative out,
// FP32 in between.  The Kerr-Schild radius is solved in FP64 from the FP64
// position (u + sqrt(u^2 + 4a^2y^2) cancels for u < 0 near the ring); the
// rest of the factored right-hand side of rhs() runs in FP32 (hardware
// reciprocal and reciprocal square root, fused multiply-adds).  Its result
// is split as dx/dl = p - corr, dp/dl = kp with corr = 2H(l.p - pt) l: the
// free-streaming part p keeps its FP64 value, and only the gravitational
// pieces corr and kp (both O(M/r)) carry FP32 rounding, so the stage
// algebra and the embedded error estimate built from these derivatives in
// FP64 see no O(1) FP32 rounding.
struct G32 {
    float a, aa, M;
    float w2c, w2f;  // Geo::w2c, Geo::w2f in FP32
};
KT_HDI G32 make_g32(const Geo& g) {
    const float aa = (float)g.aa;
    return G32{(float)g.a, aa, (float)g.M, g.exact ? 2.0f * aa : 4.0f * aa, g.exact ? 1.0f : 0.5f};
}
__device__ __forceinline__ V6 rhs32(const V6& st, float ptf, const Geo& g, const G32& gf) {
    const double x = st.x, y = st.y, z = st.z;
    // FP64 radius solve
    const double rho2 = fma(x, x, fma(y, y, z * z));
    const double u = rho2 - g.aa;
    const double ya = g.aa * y;
    const double s = sqrt_pos(fma(u, u, 4.0 * ya * y));
    const double r2d = 0.5 * (u + s);
    // FP32 bulk, in the factored form of rhs(); each reciprocal is its own
    // hardware reciprocal (a shared reciprocal of their product, as rhs()
    // forms them, costs FP32 accuracy the step control notices)
    const float xf = (float)x, yf = (float)y, zf = (float)z;
    const float pxf = (float)st.px, pyf = (float)st.py, pzf = (float)st.pz;
    const float a = gf.a, a2 = gf.aa, M = gf.M;
    const float uf = (float)u;
    const float inv_s = __builtin_amdgcn_rcpf((float)s);
    const float r2 = (float)r2d;
    const float inv_r = __builtin_amdgcn_rsqf(r2);
    const float r = r2 * inv_r;
    const float ya2 = (a2 * yf) * yf;
    const float inv_denom = __builtin_amdgcn_rcpf(r2 + a2);
    const float inv_denomH = __builtin_amdgcn_rcpf(fmaf(r2, r2, ya2));
    const float lx = fmaf(r, xf, -(a * zf)) * inv_denom;
    const float ly = yf * inv_r;
    const float lz = fmaf(r, zf, a * xf) * inv_denom;
    const float r3 = r2 * r;
    const float H2 = ((2.0f * M) * r3) * inv_denomH;  // 2 H, rounded as (2 H) would be
    const float Q = fmaf(lx, pxf, lz * pzf);
    const float lp = fmaf(ly, pyf, Q) - ptf;
    const float c = H2 * lp;
    const float inv_2r = 0.5f * inv_r;
    const float w1 = fmaf(uf, inv_s, 1.0f) * inv_2r;
    const float w2 = fmaf(uf + gf.w2c, gf.w2f * inv_s, 1.0f) * inv_2r;
    const float XP = fmaf(xf, pxf, zf * pzf);
    const float inv_r2 = inv_r * inv_r;
    const float G = fmaf(fmaf(-2.0f * r, Q, XP), inv_denom, -((yf * pyf) * inv_r2));
    const float inv_denomH2 = inv_denomH * inv_denomH;
    const float HR = ((M * r2) * inv_denomH2) * fmaf(3.0f, ya2, -(r2 * r2));
    const float HY = -((2.0f * M) * (a2 * yf)) * (r3 * inv_denomH2);
    const float lp2 = lp * lp;
    const float Z = fmaf(HR, lp2, c * G);
    const float cD = c * inv_denom;
    V6 d;
    d.x = st.px - (double)(c * lx);
    d.y = st.py - (double)(c * ly);
    d.z = st.pz - (double)(c * lz);
    d.px = (double)fmaf(xf * w1, Z, cD * fmaf(r, pxf, a * pzf));
    d.py = (double)fmaf(yf * w2, Z, fmaf(HY, lp2, (c * pyf) * inv_r));
    d.pz = (double)fmaf(zf * w1, Z, cD * fmaf(r, pzf, -(a * pxf)));
    return d;
}

// Closeness of a ray to the critical curve (the photon-orbit boundary of
// the shadow), from its flat-space constants of motion at the camera (M
// units, E = 1): J = x cross n, lambda = J_y, eta = |J|^2 - J_y^2 -
// a^2 cos^2(theta_cam).  The Boyer-Lindquist radial potential
//   R(r) = (r^2 + a^2 - a lambda)^2 - (r^2 - 2r + a^2)(eta + (lambda - a)^2)
// has a double root exactly on the critical curve; near it R has a local
// minimum close to zero between r_+ and the photon region, and the ray
// orbits ~ -log|R(r*)| times.  R is sampled on a quadratic grid in
// [r_+, 12] (dense near the horizon, where the prograde orbits of a rapidly
// spinning hole sit), the lowest sample is refined with Newton steps on
// R' = 0, and the key is -log(|R(r*)| / r*^4): large for rays that orbit
// long.  x: camera position, n: unit direction, a_m = a / M, inv_M = 1/M.
KT_HDI double critical_key(const double x[3], const double n[3], double a_m, double inv_M) {
    const double jx = x[1] * n[2] - x[2] * n[1], jy = x[2] * n[0] - x[0] * n[2], jz = x[0] * n[1] - x[1] * n[0];
    const double rc2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
    const double a = a_m;
    const double cos2 = x[1] * x[1] / rc2;
    const double lam = jy * inv_M;
    const double eta = (jx * jx + jz * jz) * (inv_M * inv_M) - a * a * cos2;
    const double rp = 1.0 + sqrt(fmax(1.0 - a * a, 0.0));
    const double c2 = eta + (lam - a) * (lam - a);
    const double r_lo = rp * 1.0001, span = 12.0 - rp;
    double best = 3.0e38, r_best = r_lo;
    const int NS = 32;
    for (int i = 0; i < NS; ++i) {
        const double t = (double)i / (double)(NS - 1);
        const double r = r_lo + span * t * t;
        const double A = r * r + a * a - a * lam;
        const double R = A * A - (r * r - 2.0 * r + a * a) * c2;
        if (R < best) {
            best = R;
            r_best = r;
        }
    }
    double r = r_best;
    for (int it = 0; it < 3; ++it) {
        const double A = r * r + a * a - a * lam;
        const double d1 = 4.0 * r * A - (2.0 * r - 2.0) * c2;
        const double d2 = 4.0 * A + 8.0 * r * r - 2.0 * c2;
        if (d2 > 0.0) r = fmin(fmax(r - d1 / d2, r_lo), 12.0);
    }
    const double A = r * r + a * a - a * lam;
    const double R = fmin(A * A - (r * r - 2.0 * r + a * a) * c2, best);
    return -log(fabs(R) / (r * r * r * r) + 1e-12);
}

// DOPRI5 continuous extension (cuda/kerr_math.cuh:186-219; Hairer, Norsett
// and Wanner's dense-output coefficients).
struct Dense {
    static constexpr double d1 = -12715105075.0 / 11282082432.0;
    static constexpr double d3 = 87487479700.0 / 32700410799.0;
    static constexpr double d4 = -10690763975.0 / 1880347072.0;
    static constexpr double d5 = 701980252875.0 / 199316789632.0;
    static constexpr double d6 = -1453857185.0 / 822651844.0;
    static constexpr double d7 = 69997945.0 / 29380423.0;
};
// d1 k1 + d3 k3 + d4 k4 + d5 k5 + d6 k6 + d7 k7 as one fma chain in stage
// order (the tracer folds it the same way while its stages run).
KT_HDI double dense_sum(double k1, double k3, double k4, double k5, double k6, double k7) {
    return fma(Dense::d7, k7, fma(Dense::d6, k6, fma(Dense::d5, k5, fma(Dense::d4, k4, fma(Dense::d3, k3, Dense::d1 * k1)))));
}
KT_HDI double dense(double y0, double y1, double k1, double k3, double k4, double k5,
                                        double k6, double k7, double h, double t) {
    double r1 = y0;
    double r2 = y1 - y0;
    double r3 = y0 + h * k1 - y1;
    double r4 = 2.0 * (y1 - y0) - h * (k1 + k7);
    double r5 = h * dense_sum(k1, k3, k4, k5, k6, k7);
    double t2 = t * t;
    double t3 = t2 * t;
    double t4 = t2 * t2;
    return r1 + t * (r2 + r3) - t2 * (r3 - r4 - r5) - t3 * (r4 + 2.0 * r5) + t4 * r5;
}
KT_HDI double dense_dt(double y0, double y1, double k1, double k3, double k4, double k5,
                                           double k6, double k7, double h, double t) {
    double r2 = y1 - y0;
    double r3 = y0 + h * k1 - y1;
    double r4 = 2.0 * (y1 - y0) - h * (k1 + k7);
    double r5 = h * dense_sum(k1, k3, k4, k5, k6, k7);
    double t2 = t * t;
    double t3 = t2 * t;
    return (r2 + r3) - 2.0 * t * (r3 - r4 - r5) - 3.0 * t2 * (r4 + 2.0 * r5) + 4.0 * t3 * r5;
}

// The same interpolant from its three step-dependent ingredients: k1, k7
// and dsum = dense_sum(k1, k3, ..., k7), so a tracer can fold the sum while
// it runs the stages instead of keeping every stage derivative.
KT_HDI double dense_s(double y0, double y1, double k1, double k7, double dsum, double h,
                                          double t) {
    double r1 = y0;
    double r2 = y1 - y0;
    double r3 = y0 + h * k1 - y1;
    double r4 = 2.0 * (y1 - y0) - h * (k1 + k7);
    double r5 = h * dsum;
    double t2 = t * t;
    double t3 = t2 * t;
    double t4 = t2 * t2;
    return r1 + t * (r2 + r3) - t2 * (r3 - r4 - r5) - t3 * (r4 + 2.0 * r5) + t4 * r5;
}
KT_HDI double dense_s_dt(double y0, double y1, double k1, double k7, double dsum, double h,
                                             double t) {
    double r2 = y1 - y0;
    double r3 = y0 + h * k1 - y1;
    double r4 = 2.0 * (y1 - y0) - h * (k1 + k7);
    double r5 = h * dsum;
    double t2 = t * t;
    double t3 = t2 * t;
    return (r2 + r3) - 2.0 * t * (r3 - r4 - r5) - 3.0 * t2 * (r4 + 2.0 * r5) + 4.0 * t3 * r5;
}
// disk_root on dense_s (identical iteration).
KT_HDI double disk_root_s(double y0, double y1, double k1, double k7, double dsum, double h) {
    double lo = 0.0, hi = 1.0, ylo = y0;
    double den = y0 - y1;
    double t = den != 0.0 ? y0 / den : 0.5;
    t = fmin(fmax(t, 0.0), 1.0);
    for (int it = 0; it < 64; ++it) {
        double yt = dense_s(y0, y1, k1, k7, dsum, h, t);
        if (yt == 0.0) return t;
        if ((yt >= 0.0) == (ylo >= 0.0)) {
            lo = t;
            ylo = yt;
        } else {
            hi = t;
        }
        double dyt = dense_s_dt(y0, y1, k1, k7, dsum, h, t);
        double tn = dyt != 0.0 ? t - yt / dyt : 0.5 * (lo + hi);
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        double step = tn - t;
        t = tn;
        if (fabs(step) <= 1e-15 || hi - lo <= 1e-15) break;
    }
    return t;
}

// Root of the interpolant y(t) on [0,1]: secant start, Newton steps kept
// inside the shrinking bracket, bisection when Newton leaves it.
KT_HDI double disk_root(double y0, double y1, double k1, double k3, double k4, double k5,
                                            double k6, double k7, double h) {
    double lo = 0.0, hi = 1.0, ylo = y0;
    double den = y0 - y1;
    double t = den != 0.0 ? y0 / den : 0.5;
    t = fmin(fmax(t, 0.0), 1.0);
    for (int it = 0; it < 64; ++it) {
        double yt = dense(y0, y1, k1, k3, k4, k5, k6, k7, h, t);
        if (yt == 0.0) return t;
        if ((yt >= 0.0) == (ylo >= 0.0)) {
            lo = t;
            ylo = yt;
        } else {
            hi = t;
        }
        double dyt = dense_dt(y0, y1, k1, k3, k4, k5, k6, k7, h, t);
        double tn = dyt != 0.0 ? t - yt / dyt : 0.5 * (lo + hi);
        if (!(tn > lo && tn < hi)) tn = 0.5 * (lo + hi);
        double step = tn - t;
        t = tn;
        if (fabs(step) <= 1e-15 || hi - lo <= 1e-15) break;
    }
    return t;
}

// cuda/kerr_math.cuh:172-179
KT_HDI double fade(double t) { return t * t * t * (t * (t * 6.0 - 15.0) + 10.0); }
KT_HDI double lerp(double a, double b, double t) { return fma(t, b - a, a); }

// cos and sin of the angle 2 pi hsh / 2^32 (a 32-bit fraction of a turn).
// The reduction is exact integer arithmetic: q = the nearest quarter turn,
// rem = hsh - q 2^30 in [-2^29, 2^29), d = rem * (2 pi 2^-32) in [-pi/4, pi/4];
// then Taylor series of sin and cos at 0 through d^17 and d^16 (the next
// terms are below 1e-19) in Horner form, and the quarter-turn rotation
// (exact swaps and negations).  Within 2 ulp of the exact values.
KT_HDI void turn_cos_sin(uint32_t hsh, double& c, double& s) {
    const uint32_t q = (hsh + 0x20000000u) >> 30;
    const int32_t rem = (int32_t)(hsh - (q << 30));
    const double d = (double)rem * 0x1.921fb54442d18p-30;
    const double z = d * d;
    double ps = fma(0x1.952c77030ad4ap-49, z, -0x1.ae7f3e733b81fp-41);
    ps = fma(ps, z, 0x1.6124613a86d09p-33);
    ps = fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = fma(ps, z, 0x1.71de3a556c734p-19);
    ps = fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = fma(ps, z, 0x1.1111111111111p-7);
    ps = fma(ps, z, -0x1.5555555555555p-3);
    const double sn = fma(d * z, ps, d);
    double pc = fma(0x1.ae7f3e733b81fp-45, z, -0x1.93974a8c07c9dp-37);
    pc = fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = fma(pc, z, 0x1.5555555555555p-5);
    pc = fma(pc, z, -0.5);
    const double cs = fma(z, pc, 1.0);
    const double a = (q & 1u) ? sn : cs;  // cos of the rotated angle, up to sign
    const double b = (q & 1u) ? cs : sn;  // sin ...
    c = ((q + 1u) & 2u) ? -a : a;         // q = 1, 2: negative cosine
    s = (q & 2u) ? -b : b;                // q = 2, 3: negative sine
}

// cuda/kerr_math.cuh:220-260: gradient noise on the integer lattice; the
// gradient angle of a lattice point comes from pcg32 of its packed
// coordinates (xor-ed with the seed mix, 0 = reference hash) as a fraction
// hsh / 2^32 of a turn.
KT_HDI void lattice_gradient(int ix, int iy, uint64_t seed_mix, double& gx, double& gy) {
    uint64_t st = ((uint64_t)(uint32_t)ix << 32) | (uint64_t)(uint32_t)iy;
    st ^= seed_mix;
    turn_cos_sin(pcg32(st), gx, gy);
}
KT_HDI double gradient_noise(double u, double v, uint64_t seed_mix) {
    double fu = floor(u), fv = floor(v);
    int ix0 = (int)fu, iy0 = (int)fv;
    double fx = u - (double)ix0;
    double fy = v - (double)iy0;
    double g00x, g00y, g10x, g10y, g01x, g01y, g11x, g11y;
    lattice_gradient(ix0, iy0, seed_mix, g00x, g00y);
    lattice_gradient(ix0 + 1, iy0, seed_mix, g10x, g10y);
    lattice_gradient(ix0, iy0 + 1, seed_mix, g01x, g01y);
    lattice_gradient(ix0 + 1, iy0 + 1, seed_mix, g11x, g11y);
    double n00 = g00x * fx + g00y * fy;
    double n10 = g10x * (fx - 1.0) + g10y * fy;
    double n01 = g01x * fx + g01y * (fy - 1.0);
    double n11 = g11x * (fx - 1.0) + g11y * (fy - 1.0);
    double u_f = fade(fx);
    double v_f = fade(fy);
    double x1 = lerp(n00, n10, u_f);
    double x2 = lerp(n01, n11, u_f);
    return lerp(x1, x2, v_f);
}
// cuda/kerr_math.cuh:261-281
KT_HDI double fbm(double u, double v, int octaves, uint64_t seed_mix) {
    if (octaves <= 0) return 0.5;
    double sum = 0.0, amp = 1.0, freq = 1.0, amp_sum = 0.0;
    for (int i = 0; i < octaves; ++i) {
        sum = sum + gradient_noise(u * freq, v * freq, seed_mix) * amp;
        amp_sum += amp;
        amp *= 0.5;
        freq *= 2.0;
    }
    double vv = amp_sum > 0.0 ? (sum / amp_sum) : 0.0;
    return vv * 0.5 + 0.5;
}

}  // namespace kt
