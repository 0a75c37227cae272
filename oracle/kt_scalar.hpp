// Scalar types for the FP64 oracle: plain double, a forward-mode dual
// number with N tangent directions, and an operation-counting scalar.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product links this file.
//
// Every routine of the oracle is a template over one of these types so that
// the same restated source gives (1) the FP64 reference render, (2) the
// forward-mode gradient of the render with respect to spin, inclination,
// observer distance and accretion-rate scale, and (3) exact algorithmic
// FLOP counts of the right-hand side and of one DP5(4) attempt.
#pragma once
#include <cmath>
#include <cstdint>
#include <limits>

namespace kto {

// ----------------------------------------------------------------- double
inline double val(double x) { return x; }
inline double fma_(double a, double b, double c) { return std::fma(a, b, c); }
inline double sqrt_(double x) { return std::sqrt(x); }
inline double exp_(double x) { return std::exp(x); }
inline double log_(double x) { return std::log(x); }
inline double sin_(double x) { return std::sin(x); }
inline double cos_(double x) { return std::cos(x); }
inline double acos_(double x) { return std::acos(x); }
inline double asin_(double x) { return std::asin(x); }
inline double cbrt_(double x) { return std::cbrt(x); }
inline double atan2_(double y, double x) { return std::atan2(y, x); }
inline double fabs_(double x) { return std::fabs(x); }
// C99 fmax/fmin: a NaN operand yields the other operand, as CUDA's fmaxf.
inline double fmax_(double a, double b) { return std::fmax(a, b); }
inline double fmin_(double a, double b) { return std::fmin(a, b); }
inline double fmod_(double a, double b) { return std::fmod(a, b); }

// x^(-1/5) for the step-size controller (cuda/kerr_trace.cuh:119 uses
// __powf(error, -0.2f)).  Written from the mathematics so that the CPU and
// the GPU evaluate the same operation sequence (only IEEE-exact +,-,*,/,fma
// and power-of-two scaling):
//   x = m * 2^e with m in [1,2);  e = 5q + j, j in 0..4
//   x^(-1/5) = m^(-1/5) * 2^(-j/5) * 2^(-q)
// m^(-1/5) starts from a quadratic fit and takes four Newton steps on
// f(y) = y^-5 - m, i.e. y <- y * (6 - m y^5) / 5 (relative error
// e -> 3e^2: 2e-3 -> 1e-5 -> 4e-10 -> 5e-19 -> rounding).  2^(-j/5) are correctly rounded
// constants (50-digit values, rounded once to binary64).
inline double pow_m02(double x) {
    if (std::isnan(x)) return x;
    if (x == 0.0) return std::numeric_limits<double>::infinity();
    if (std::isinf(x)) return x > 0 ? 0.0 : std::numeric_limits<double>::quiet_NaN();
    if (x < 0.0) return std::numeric_limits<double>::quiet_NaN();
    int e = 0;
    double m = std::frexp(x, &e);  // m in [0.5, 1)
    m = m * 2.0;                   // m in [1, 2)
    e = e - 1;
    int q = e >= 0 ? e / 5 : -((-e + 4) / 5);
    int j = e - 5 * q;             // 0..4
    static const double two_m_j5[5] = {
        1.0,
        0x1.bdb8cdadbe120p-1,  // 2^(-1/5)
        0x1.8406003b2ae5cp-1,  // 2^(-2/5)
        0x1.51cb453b9536cp-1,  // 2^(-3/5)
        0x1.2611186bae675p-1,  // 2^(-4/5)
    };
    // quadratic least-squares start for m^(-1/5) on [1,2), |error| < 2e-3
    double y = std::fma(std::fma(0.0517578125, m, -0.283203125), m, 1.2294921875);
    for (int it = 0; it < 4; ++it) {
        double y2 = y * y;
        double y5 = y2 * y2 * y;
        y = y * (6.0 - m * y5) * 0.2;
    }
    return std::ldexp(y * two_m_j5[j], -q);
}

// cos and sin of 2 pi h / 2^32 (h a 32-bit fraction of a turn): exact
// integer reduction to the nearest quarter turn, d = rem * (2 pi 2^-32) in
// [-pi/4, pi/4], Taylor series of sin through d^17 and of cos through d^16
// (remainders < 1e-19), exact quarter-turn rotation.
inline void turn_cos_sin(uint32_t h, double& c, double& s) {
    const uint32_t q = (h + 0x20000000u) >> 30;
    const int32_t rem = (int32_t)(h - (q << 30));
    const double d = (double)rem * 0x1.921fb54442d18p-30;  // 2 pi / 2^32
    const double z = d * d;
    // (-1)^k / (2k+1)!, k = 8..1
    double ps = std::fma(0x1.952c77030ad4ap-49, z, -0x1.ae7f3e733b81fp-41);
    ps = std::fma(ps, z, 0x1.6124613a86d09p-33);
    ps = std::fma(ps, z, -0x1.ae64567f544e4p-26);
    ps = std::fma(ps, z, 0x1.71de3a556c734p-19);
    ps = std::fma(ps, z, -0x1.a01a01a01a01ap-13);
    ps = std::fma(ps, z, 0x1.1111111111111p-7);
    ps = std::fma(ps, z, -0x1.5555555555555p-3);
    const double sn = std::fma(d * z, ps, d);
    // (-1)^k / (2k)!, k = 8..1
    double pc = std::fma(0x1.ae7f3e733b81fp-45, z, -0x1.93974a8c07c9dp-37);
    pc = std::fma(pc, z, 0x1.1eed8eff8d898p-29);
    pc = std::fma(pc, z, -0x1.27e4fb7789f5cp-22);
    pc = std::fma(pc, z, 0x1.a01a01a01a01ap-16);
    pc = std::fma(pc, z, -0x1.6c16c16c16c17p-10);
    pc = std::fma(pc, z, 0x1.5555555555555p-5);
    pc = std::fma(pc, z, -0.5);
    const double cs = std::fma(z, pc, 1.0);
    const double a = (q & 1u) ? sn : cs;
    const double b = (q & 1u) ? cs : sn;
    c = ((q + 1u) & 2u) ? -a : a;
    s = (q & 2u) ? -b : b;
}

// ------------------------------------------------------------ dual number
template <int N>
struct Dual {
    double v;
    double d[N];
    Dual() : v(0.0) { for (int i = 0; i < N; ++i) d[i] = 0.0; }
    Dual(double x) : v(x) { for (int i = 0; i < N; ++i) d[i] = 0.0; }  // NOLINT
    static Dual seed(double x, int k) { Dual r(x); r.d[k] = 1.0; return r; }
};
template <int N> inline double val(const Dual<N>& x) { return x.v; }
template <int N> inline Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r; r.v = -a.v; for (int i = 0; i < N; ++i) r.d[i] = -a.d[i]; return r;
}
template <int N> inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v + b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] + b.d[i]; return r;
}
template <int N> inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v - b.v; for (int i = 0; i < N; ++i) r.d[i] = a.d[i] - b.d[i]; return r;
}
template <int N> inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
template <int N> inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r; r.v = a.v / b.v;
    for (int i = 0; i < N; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) / b.v;
    return r;
}
template <int N> inline Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> inline Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> inline Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> inline Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N> inline Dual<N> operator*(const Dual<N>& a, double b) { return a * Dual<N>(b); }
template <int N> inline Dual<N> operator*(double a, const Dual<N>& b) { return Dual<N>(a) * b; }
template <int N> inline Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
template <int N> inline Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }
template <int N> inline Dual<N>& operator+=(Dual<N>& a, const Dual<N>& b) { a = a + b; return a; }
template <int N> inline Dual<N>& operator*=(Dual<N>& a, const Dual<N>& b) { a = a * b; return a; }
#define KTO_DUAL_CMP(OP) \
    template <int N> inline bool operator OP(const Dual<N>& a, const Dual<N>& b) { return a.v OP b.v; } \
    template <int N> inline bool operator OP(const Dual<N>& a, double b) { return a.v OP b; } \
    template <int N> inline bool operator OP(double a, const Dual<N>& b) { return a OP b.v; }
KTO_DUAL_CMP(<) KTO_DUAL_CMP(>) KTO_DUAL_CMP(<=) KTO_DUAL_CMP(>=) KTO_DUAL_CMP(==) KTO_DUAL_CMP(!=)
#undef KTO_DUAL_CMP
template <int N> inline Dual<N> chain(const Dual<N>& a, double fv, double dfdx) {
    Dual<N> r; r.v = fv; for (int i = 0; i < N; ++i) r.d[i] = dfdx * a.d[i]; return r;
}
template <int N> inline Dual<N> fma_(const Dual<N>& a, const Dual<N>& b, const Dual<N>& c) {
    Dual<N> r; r.v = std::fma(a.v, b.v, c.v);
    for (int i = 0; i < N; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i] + c.d[i];
    return r;
}
template <int N> inline Dual<N> sqrt_(const Dual<N>& a) { double s = std::sqrt(a.v); return chain(a, s, 0.5 / s); }
template <int N> inline Dual<N> exp_(const Dual<N>& a) { double e = std::exp(a.v); return chain(a, e, e); }
template <int N> inline Dual<N> log_(const Dual<N>& a) { return chain(a, std::log(a.v), 1.0 / a.v); }
template <int N> inline Dual<N> sin_(const Dual<N>& a) { return chain(a, std::sin(a.v), std::cos(a.v)); }
template <int N> inline Dual<N> cos_(const Dual<N>& a) { return chain(a, std::cos(a.v), -std::sin(a.v)); }
template <int N> inline Dual<N> acos_(const Dual<N>& a) { return chain(a, std::acos(a.v), -1.0 / std::sqrt(1.0 - a.v * a.v)); }
template <int N> inline Dual<N> asin_(const Dual<N>& a) { return chain(a, std::asin(a.v), 1.0 / std::sqrt(1.0 - a.v * a.v)); }
template <int N> inline Dual<N> cbrt_(const Dual<N>& a) { double c = std::cbrt(a.v); return chain(a, c, 1.0 / (3.0 * c * c)); }
template <int N> inline Dual<N> atan2_(const Dual<N>& y, const Dual<N>& x) {
    Dual<N> r; r.v = std::atan2(y.v, x.v); double q = x.v * x.v + y.v * y.v;
    for (int i = 0; i < N; ++i) r.d[i] = (x.v * y.d[i] - y.v * x.d[i]) / q;
    return r;
}
template <int N> inline Dual<N> fabs_(const Dual<N>& a) { return a.v < 0 ? -a : a; }
template <int N> inline Dual<N> fmax_(const Dual<N>& a, const Dual<N>& b) {
    if (std::isnan(a.v)) return b;
    if (std::isnan(b.v)) return a;
    return a.v >= b.v ? a : b;
}
template <int N> inline Dual<N> fmin_(const Dual<N>& a, const Dual<N>& b) {
    if (std::isnan(a.v)) return b;
    if (std::isnan(b.v)) return a;
    return a.v <= b.v ? a : b;
}
template <int N> inline Dual<N> fmax_(const Dual<N>& a, double b) { return fmax_(a, Dual<N>(b)); }
template <int N> inline Dual<N> fmin_(const Dual<N>& a, double b) { return fmin_(a, Dual<N>(b)); }
template <int N> inline Dual<N> fmod_(const Dual<N>& a, double b) { Dual<N> r = a; r.v = std::fmod(a.v, b); return r; }
template <int N> inline Dual<N> pow_m02(const Dual<N>& a) { double p = pow_m02(a.v); return chain(a, p, -0.2 * p / a.v); }

// ------------------------------------------------- operation-counting type
// Counts IEEE floating-point operations: add, sub, mul, div, sqrt count 1,
// fma counts 2; comparisons, min/max/abs and conversions count 0;
// transcendental library calls are tallied separately.
struct OpCount {
    uint64_t flops = 0;
    uint64_t transcendental = 0;
};
inline OpCount& op_counter() { static thread_local OpCount c; return c; }

struct Fl {
    double v;
    Fl() : v(0.0) {}
    Fl(double x) : v(x) {}  // NOLINT
};
inline double val(const Fl& x) { return x.v; }
inline Fl fl_op(double r, uint64_t n) { op_counter().flops += n; return Fl(r); }
inline Fl operator-(const Fl& a) { return Fl(-a.v); }
inline Fl operator+(const Fl& a, const Fl& b) { return fl_op(a.v + b.v, 1); }
inline Fl operator-(const Fl& a, const Fl& b) { return fl_op(a.v - b.v, 1); }
inline Fl operator*(const Fl& a, const Fl& b) { return fl_op(a.v * b.v, 1); }
inline Fl operator/(const Fl& a, const Fl& b) { return fl_op(a.v / b.v, 1); }
inline Fl operator+(const Fl& a, double b) { return a + Fl(b); }
inline Fl operator+(double a, const Fl& b) { return Fl(a) + b; }
inline Fl operator-(const Fl& a, double b) { return a - Fl(b); }
inline Fl operator-(double a, const Fl& b) { return Fl(a) - b; }
inline Fl operator*(const Fl& a, double b) { return a * Fl(b); }
inline Fl operator*(double a, const Fl& b) { return Fl(a) * b; }
inline Fl operator/(const Fl& a, double b) { return a / Fl(b); }
inline Fl operator/(double a, const Fl& b) { return Fl(a) / b; }
inline Fl& operator+=(Fl& a, const Fl& b) { a = a + b; return a; }
inline Fl& operator*=(Fl& a, const Fl& b) { a = a * b; return a; }
#define KTO_FL_CMP(OP) \
    inline bool operator OP(const Fl& a, const Fl& b) { return a.v OP b.v; } \
    inline bool operator OP(const Fl& a, double b) { return a.v OP b; } \
    inline bool operator OP(double a, const Fl& b) { return a OP b.v; }
KTO_FL_CMP(<) KTO_FL_CMP(>) KTO_FL_CMP(<=) KTO_FL_CMP(>=) KTO_FL_CMP(==) KTO_FL_CMP(!=)
#undef KTO_FL_CMP
inline Fl fma_(const Fl& a, const Fl& b, const Fl& c) { return fl_op(std::fma(a.v, b.v, c.v), 2); }
inline Fl sqrt_(const Fl& a) { return fl_op(std::sqrt(a.v), 1); }
inline Fl tr_op(double r) { op_counter().transcendental += 1; return Fl(r); }
inline Fl exp_(const Fl& a) { return tr_op(std::exp(a.v)); }
inline Fl log_(const Fl& a) { return tr_op(std::log(a.v)); }
inline Fl sin_(const Fl& a) { return tr_op(std::sin(a.v)); }
inline Fl cos_(const Fl& a) { return tr_op(std::cos(a.v)); }
inline Fl acos_(const Fl& a) { return tr_op(std::acos(a.v)); }
inline Fl asin_(const Fl& a) { return tr_op(std::asin(a.v)); }
inline Fl cbrt_(const Fl& a) { return tr_op(std::cbrt(a.v)); }
inline Fl atan2_(const Fl& y, const Fl& x) { return tr_op(std::atan2(y.v, x.v)); }
inline Fl fabs_(const Fl& a) { return Fl(std::fabs(a.v)); }
inline Fl fmax_(const Fl& a, const Fl& b) { return Fl(std::fmax(a.v, b.v)); }
inline Fl fmin_(const Fl& a, const Fl& b) { return Fl(std::fmin(a.v, b.v)); }
inline Fl fmod_(const Fl& a, double b) { return tr_op(std::fmod(a.v, b)); }
inline Fl pow_m02(const Fl& a) { return tr_op(pow_m02(a.v)); }

}  // namespace kto
