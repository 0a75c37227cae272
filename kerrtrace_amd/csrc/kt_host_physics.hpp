// Host-side disk physics shared by the renderer (double) and the
// differentiable renderer (a forward-mode dual number carrying d/d spin):
// the prograde ISCO and the Novikov-Thorne flux factor as the reference
// writes them (src/renderer/utils.rs:11-52).  The double instantiation is
// the same operation sequence as a plain double function.
#pragma once
#include <cmath>

namespace kt_host {

static const double kPiH = 3.14159265358979323846;

// value + one tangent
struct HD {
    double v, d;
};
inline double val(double x) { return x; }
inline double val(const HD& x) { return x.v; }
inline HD operator+(HD a, HD b) { return HD{a.v + b.v, a.d + b.d}; }
inline HD operator-(HD a, HD b) { return HD{a.v - b.v, a.d - b.d}; }
inline HD operator-(HD a) { return HD{-a.v, -a.d}; }
inline HD operator*(HD a, HD b) { return HD{a.v * b.v, a.d * b.v + a.v * b.d}; }
inline HD operator/(HD a, HD b) {
    double q = a.v / b.v;
    return HD{q, (a.d - q * b.d) / b.v};
}
inline HD operator+(HD a, double b) { return HD{a.v + b, a.d}; }
inline HD operator+(double a, HD b) { return HD{a + b.v, b.d}; }
inline HD operator-(HD a, double b) { return HD{a.v - b, a.d}; }
inline HD operator-(double a, HD b) { return HD{a - b.v, -b.d}; }
inline HD operator*(HD a, double b) { return HD{a.v * b, a.d * b}; }
inline HD operator*(double a, HD b) { return HD{a * b.v, a * b.d}; }
inline HD operator/(HD a, double b) { return HD{a.v / b, a.d / b}; }
inline HD operator/(double a, HD b) {
    double q = a / b.v;
    return HD{q, -q * b.d / b.v};
}

inline double m_sqrt(double x) { return std::sqrt(x); }
inline double m_cbrt(double x) { return std::cbrt(x); }
inline double m_acos(double x) { return std::acos(x); }
inline double m_cos(double x) { return std::cos(x); }
inline double m_log(double x) { return std::log(x); }
inline double m_fma(double a, double b, double c) { return std::fma(a, b, c); }
inline double m_fmax0(double x) { return std::fmax(x, 0.0); }
inline HD m_sqrt(HD a) {
    double r = std::sqrt(a.v);
    return HD{r, a.d * 0.5 / r};
}
inline HD m_cbrt(HD a) {
    double c = std::cbrt(a.v);
    return HD{c, a.d / (3.0 * c * c)};
}
inline HD m_acos(HD a) { return HD{std::acos(a.v), -a.d / std::sqrt(1.0 - a.v * a.v)}; }
inline HD m_cos(HD a) { return HD{std::cos(a.v), -std::sin(a.v) * a.d}; }
inline HD m_log(HD a) { return HD{std::log(a.v), a.d / a.v}; }
inline HD m_fma(HD a, HD b, HD c) { return HD{std::fma(a.v, b.v, c.v), a.d * b.v + a.v * b.d + c.d}; }
inline HD m_fmax0(HD a) { return a.v >= 0.0 ? a : HD{0.0, 0.0}; }

// Prograde/retrograde ISCO (utils.rs:11-20).
template <class S>
S calc_isco(S a_norm, bool prograde, double mass) {
    S aa = a_norm * a_norm;
    S z1 = m_fma(m_cbrt(1.0 - aa), m_cbrt(1.0 + a_norm) + m_cbrt(1.0 - a_norm), S{1.0});
    S z2 = m_sqrt(m_fma(S{3.0}, aa, z1 * z1));
    double sign = prograde ? -1.0 : 1.0;
    S term_inside = (3.0 - z1) * m_fma(S{2.0}, z2, 3.0 + z1);
    return mass * (3.0 + z2 + sign * m_sqrt(m_fmax0(term_inside)));
}

// Novikov-Thorne flux factor exactly as the reference writes it
// (utils.rs:21-52): of the three cubic-root log terms only those whose
// root denominator is positive are kept.
template <class S>
S nt_factor(S r, S a_norm, S r_isco, double inv_m) {
    if (val(r) <= val(r_isco)) return S{0.0};
    S r_norm = r * inv_m;
    S x = m_sqrt(r_norm);
    S x_ms = m_sqrt(r_isco * inv_m);
    S base = m_acos(-a_norm) * (1.0 / 3.0);
    const double step = 2.0 * kPiH / 3.0;
    const S roots[3] = {2.0 * m_cos(base), 2.0 * m_cos(base - step), 2.0 * m_cos(base + step)};
    S sum_log{0.0};
    for (int i = 0; i < 3; ++i) {
        S xi = roots[i];
        S denom = xi * (xi - roots[(i + 1) % 3]) * (xi - roots[(i + 2) % 3]);
        if (val(denom) <= 0.0) continue;
        S coef = 3.0 * (xi - a_norm) * (xi - a_norm) / denom;
        S v = (x - xi) / (x_ms - xi);
        if (val(v) > 0.0) sum_log = sum_log + coef * m_log(v);
    }
    S q = m_fma(1.5 * a_norm, -m_log(x / x_ms), x - x_ms) - sum_log;
    S geo = r_norm * m_fma(S{2.0}, a_norm, m_fma(r_norm, x, -(3.0 * x)));
    return m_fmax0(q / geo);
}

}  // namespace kt_host
