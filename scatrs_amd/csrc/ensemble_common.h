// Shared pieces of the ensemble path: the distributions' densities and radius ranges, and the
// per-ensemble input record the kernels read.
#pragma once

#include <cmath>

#include "cx.h"

namespace scatrs {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kMaxLayers = 16;

// The radius where the unnormalized density is scaled near 1 (distribution.rs:293-309).
SCATRS_HD double dist_reference(int kind, double p0, double p1) {
    if (kind == 0) return p0;
    const double alpha = (1.0 - 3.0 * p1) / p1;
    return alpha > 0.0 ? alpha * p0 * p1 : p0;
}

// Whether the distribution's parameters are valid (distribution.rs:99-185).
SCATRS_HD bool dist_valid(int kind, double p0, double p1) {
    if (kind == 0) return __builtin_isfinite(p0) && p0 > 0.0 && __builtin_isfinite(p1) && p1 > 1.0;
    if (kind == 1) return __builtin_isfinite(p0) && p0 > 0.0 && p1 > 0.0 && p1 < 0.5;
    return false;
}

// Upper end of the gamma range: r_hi = a (1 + G b + sqrt(2 G b)). The area-weighted gamma law is
// Gamma(1/b, scale a b) in r, whose upper tail beyond k + c (k = 1/b, in units of the scale) is
// at most exp(-(c - k ln(1 + c/k))) (Chernoff); c = G + sqrt(2 k G) keeps that below e^-G.
constexpr double kGammaTail = 37.0;

// The fixed ln r range: lognormal ln r_g -+ 5 ln sigma_g; gamma from ln a - 8 s, with
// s = sqrt(b / (1 - 2 b)) the reference's log-width of the gamma peak (distribution.rs:319-323),
// to ln a + ln(1 + G b + sqrt(2 G b)), G = kGammaTail.
SCATRS_HD void radius_range(int kind, double p0, double p1, double *t_lo, double *t_hi) {
    if (kind == 0) {
        const double c = log(p0), h = 5.0 * log(p1);
        *t_lo = c - h;
        *t_hi = c + h;
    } else {
        const double s = sqrt(p1 / (1.0 - 2.0 * p1));
        const double c = log(p0);
        *t_lo = c - 8.0 * s;
        *t_hi = c + log(1.0 + kGammaTail * p1 + sqrt(2.0 * kGammaTail * p1));
    }
}

// Derivatives of the range's midpoint and half-width in ln r with respect to (p0, p1), and the
// half-width itself.
SCATRS_HD double range_derivatives(int kind, double p0, double p1, double dmid[2],
                                   double dhalf[2]) {
    if (kind == 0) {
        dmid[0] = 1.0 / p0;
        dmid[1] = 0.0;
        dhalf[0] = 0.0;
        dhalf[1] = 5.0 / p1;
        return 5.0 * log(p1);
    }
    const double q = 1.0 - 2.0 * p1;
    const double s = sqrt(p1 / q);
    const double ds = 0.5 / (s * q * q);  // d s / d b
    const double rt = sqrt(2.0 * kGammaTail * p1);
    const double up = log(1.0 + kGammaTail * p1 + rt);  // t_hi - ln a
    const double dup = (kGammaTail + kGammaTail / rt) / (1.0 + kGammaTail * p1 + rt);
    dmid[0] = 1.0 / p0;
    dmid[1] = 0.5 * (dup - 8.0 * ds);
    dhalf[0] = 0.0;
    dhalf[1] = 0.5 * (dup + 8.0 * ds);
    return 0.5 * (up + 8.0 * s);
}

// d(ln u)/d(ln r) at fixed parameters.
SCATRS_HD double log_density_slope(int kind, double p0, double p1, double r) {
    if (kind == 0) {
        const double s = log(p1);
        return -(log(r) - log(p0)) / (s * s) - 1.0;
    }
    const double alpha = (1.0 - 3.0 * p1) / p1;
    return alpha - r / (p0 * p1);
}

// ln u(r) and d(ln u)/d(p0, p1), with the reference radius held fixed
// (distribution.rs:366-430).
SCATRS_HD double log_density(int kind, double p0, double p1, double r, double *d0, double *d1) {
    const double r0 = dist_reference(kind, p0, p1);
    if (kind == 0) {
        const double mu = log(p0), s = log(p1);
        const double z = log(r) - mu;
        if (d0) *d0 = z / (s * s * p0);
        if (d1) *d1 = (z * z / (s * s * s) - 1.0 / s) / p1;
        return -z * z / (2.0 * s * s) - log(s) - log(r / r0);
    }
    const double a = p0, b = p1;
    const double alpha = (1.0 - 3.0 * b) / b;
    const double scale = a * b;
    const double l = log(r / r0);
    if (d0) *d0 = (r - r0) * b / (scale * scale);
    if (d1) *d1 = -l / (b * b) + (r - r0) * a / (scale * scale);
    return alpha * l - (r - r0) / scale;
}

}  // namespace scatrs
