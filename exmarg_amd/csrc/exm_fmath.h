// Float64 logarithm and arc tangent for results rounded to float32: the matched filter's ln|z| and
// arg z.  Host and device, so that the same code is checked against the C library on the CPU.
#pragma once

#include <cmath>

#include <hip/hip_runtime.h>

namespace exm {

// ln p for finite p > 0, for outputs rounded to float32 (relative error ~3e-16): p = m 2^e with m in
// [1/sqrt 2, sqrt 2), ln m = 2 atanh(s) = 2 s (1 + z/3 + z^2/5 + ...), s = (m - 1) / (m + 1), z = s^2
// <= 0.0295, nine terms (the first omitted z^9 / 19 < 1e-15)
__host__ __device__ inline double ln_pos(double p) {
    int e;
    double m = std::frexp(p, &e);
    const bool low = m < 0.70710678118654752440;
    m = low ? 2.0 * m : m;
    e = low ? e - 1 : e;
    const double f = m - 1.0;
    const double s = f / (2.0 + f);
    const double z = s * s;
    double q = 1.0 / 17.0;
    q = std::fma(q, z, 1.0 / 15.0);
    q = std::fma(q, z, 1.0 / 13.0);
    q = std::fma(q, z, 1.0 / 11.0);
    q = std::fma(q, z, 1.0 / 9.0);
    q = std::fma(q, z, 1.0 / 7.0);
    q = std::fma(q, z, 1.0 / 5.0);
    q = std::fma(q, z, 1.0 / 3.0);
    const double s2 = s + s;
    return std::fma((double)e, 0.69314718055994530942, std::fma(s2 * z, q, s2));
}

// atan2(y, x) for outputs rounded to float32 (error ~2e-16 of the result): the angle from the nearer
// axis, a = atan(mn / mx) in [0, pi/4], is atan(t) with t = mn / mx up to tan(pi/8), else pi/4 +
// atan(t) with t = (mn - mx) / (mn + mx) (one division either way), |t| <= tan(pi/8) = 0.414, by 19 terms
// of t (1 - z/3 + z^2/5 - ...) (the first omitted t z^19 / 39 < 1e-16 t); then the octant.
// atan2(+-0, x) is +-0 for x >= 0 and +-pi for x < 0.
__host__ __device__ inline double atan2_f(double y, double x) {
    const double ax = std::fabs(x), ay = std::fabs(y);
    const double mx = std::fmax(ax, ay), mn = std::fmin(ax, ay);
    const bool mid = mn > 0.41421356237309504880 * mx;
    const double num = mid ? mn - mx : mn;
    const double den = mid ? mn + mx : (mx > 0.0 ? mx : 1.0);
    const double t = num / den;
    const double z = t * t;
    double q = 1.0 / 39.0;
    q = std::fma(q, -z, 1.0 / 37.0);
    q = std::fma(q, -z, 1.0 / 35.0);
    q = std::fma(q, -z, 1.0 / 33.0);
    q = std::fma(q, -z, 1.0 / 31.0);
    q = std::fma(q, -z, 1.0 / 29.0);
    q = std::fma(q, -z, 1.0 / 27.0);
    q = std::fma(q, -z, 1.0 / 25.0);
    q = std::fma(q, -z, 1.0 / 23.0);
    q = std::fma(q, -z, 1.0 / 21.0);
    q = std::fma(q, -z, 1.0 / 19.0);
    q = std::fma(q, -z, 1.0 / 17.0);
    q = std::fma(q, -z, 1.0 / 15.0);
    q = std::fma(q, -z, 1.0 / 13.0);
    q = std::fma(q, -z, 1.0 / 11.0);
    q = std::fma(q, -z, 1.0 / 9.0);
    q = std::fma(q, -z, 1.0 / 7.0);
    q = std::fma(q, -z, 1.0 / 5.0);
    q = std::fma(q, -z, 1.0 / 3.0);
    double a = std::fma(-t * z, q, t);
    a = mid ? a + 0.78539816339744830962 : a;
    a = ay > ax ? 1.57079632679489661923 - a : a;
    a = x < 0.0 ? 3.14159265358979323846 - a : a;
    return std::copysign(a, y);
}

}  // namespace exm
