// Shared pieces of the ensemble path: the distributions' densities and radius ranges, and the
// per-ensemble input record the kernels read.
//
// Distributions (scatrs/src/ensemble/distribution.rs:15-68) and their parameters p[3]:
//   0 lognormal (r_g, sigma_g, -), 1 gamma (a = r_eff, b = v_eff, -),
//   2 modified gamma (alpha, b, gamma), 3 power law (alpha; r_min, r_max, which are fixed),
//   4 monodisperse (radius, -, -), 5 one segment of a tabulated law (r_0, r_1, tau): n linear
//   on [r_0, r_1], n ~ (1 - tau)(1 - s) + tau s with s = (r - r_0)/(r_1 - r_0), no parameters
//   to differentiate (a tabulated law is the mixture of its segments, each weighted by its
//   number integral; distribution.rs:55-61,432-441).
#pragma once

#include <cmath>

#include "cx.h"

namespace scatrs {

constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr int kMaxLayers = 16;

enum : int {
    kLogNormal = 0,
    kGamma = 1,
    kModifiedGamma = 2,
    kPowerLaw = 3,
    kMonodisperse = 4,
    kSegment = 5
};

// The radius where the unnormalized density is scaled near 1 (distribution.rs:293-318).
SCATRS_HD double dist_reference(int kind, const double *p) {
    if (kind == kLogNormal) return p[0];
    if (kind == kGamma) {
        const double alpha = (1.0 - 3.0 * p[1]) / p[1];
        return alpha > 0.0 ? alpha * p[0] * p[1] : p[0];
    }
    if (kind == kModifiedGamma) return pow(fmax(p[0], 1.0) / (p[1] * p[2]), 1.0 / p[2]);
    if (kind == kPowerLaw) return p[0] >= 0.0 ? p[1] : p[2];
    return 1.0;
}

// Whether the distribution's parameters are valid (distribution.rs:99-185).
SCATRS_HD bool dist_valid(int kind, const double *p) {
    const bool f0 = __builtin_isfinite(p[0]), f1 = __builtin_isfinite(p[1]),
               f2 = __builtin_isfinite(p[2]);
    if (kind == kLogNormal) return f0 && p[0] > 0.0 && f1 && p[1] > 1.0;
    if (kind == kGamma) return f0 && p[0] > 0.0 && p[1] > 0.0 && p[1] < 0.5;
    if (kind == kModifiedGamma) return f0 && p[0] > -1.0 && f1 && p[1] > 0.0 && f2 && p[2] > 0.0;
    if (kind == kPowerLaw) return f0 && f1 && p[1] > 0.0 && f2 && p[2] > p[1];
    if (kind == kMonodisperse) return f0 && p[0] > 0.0;
    if (kind == kSegment) return f0 && p[0] > 0.0 && f1 && p[1] > p[0] && p[2] >= 0.0 && p[2] <= 1.0;
    return false;
}

// Upper end of the gamma range: r_hi = a (1 + G b + sqrt(2 G b)). The area-weighted gamma law is
// Gamma(1/b, scale a b) in r, whose upper tail beyond k + c (k = 1/b, in units of the scale) is
// at most exp(-(c - k ln(1 + c/k))) (Chernoff); c = G + sqrt(2 k G) keeps that below e^-G.
constexpr double kGammaTail = 37.0;

// psi(x), x > 0: the recurrence psi(x) = psi(x + 1) - 1/x up to x >= 6, then the asymptotic
// series ln x - 1/(2x) - sum_k B_2k / (2k x^2k) (Abramowitz & Stegun 6.3.18).
SCATRS_HD double digamma(double x) {
    double acc = 0.0;
    while (x < 6.0) {
        acc -= 1.0 / x;
        x += 1.0;
    }
    const double q = 1.0 / (x * x);
    const double series =
        q * (1.0 / 12.0 - q * (1.0 / 120.0 - q * (1.0 / 252.0 - q * (1.0 / 240.0 - q * (1.0 / 132.0)))));
    return acc + log(x) - 0.5 / x - series;
}

// The modified gamma law in y = b r^gamma: its number density is Gamma((alpha + 1)/gamma) in
// y, its area-weighted one Gamma(kA = (alpha + 3)/gamma). The range bounds both tails of the
// area-weighted law (the cross sections'): below, P(Y < y) <= y^kA / Gamma(kA + 1) = e^-G,
// ln y_lo = (ln Gamma(kA + 1) - G)/kA; above, as for the gamma law, y_hi = kA + G + sqrt(2 G kA);
// t = (ln y - ln b)/gamma. Number averages are those of the law truncated to the range. With
// dlo, dhi not null, the derivatives of (t_lo, t_hi) with respect to the parameters.
SCATRS_HD void mgamma_range(const double *p, double *t_lo, double *t_hi, double *dlo,
                            double *dhi) {
    const double G = kGammaTail;
    const double alpha = p[0], b = p[1], g = p[2];
    const double ka = (alpha + 3.0) / g;
    const double lylo = (lgamma(ka + 1.0) - G) / ka;
    const double rt = sqrt(2.0 * G * ka);
    const double yhi = ka + G + rt;
    *t_lo = (lylo - log(b)) / g;
    *t_hi = (log(yhi) - log(b)) / g;
    if (dlo) {
        const double dlylo_dka = (digamma(ka + 1.0) - lylo) / ka;
        const double dlyhi_dka = (1.0 + G / rt) / yhi;
        // dkA/dalpha = 1/gamma, dkA/dgamma = -kA/gamma
        dlo[0] = dlylo_dka / (g * g);
        dlo[1] = -1.0 / (b * g);
        dlo[2] = dlylo_dka * (-ka / g) / g - *t_lo / g;
        dhi[0] = dlyhi_dka / (g * g);
        dhi[1] = -1.0 / (b * g);
        dhi[2] = dlyhi_dka * (-ka / g) / g - *t_hi / g;
    }
}

// The fixed ln r range: lognormal ln r_g -+ 5 ln sigma_g; gamma from ln a - 8 s, with
// s = sqrt(b / (1 - 2 b)) the reference's log-width of the gamma peak (distribution.rs:319-323),
// to ln a + ln(1 + G b + sqrt(2 G b)), G = kGammaTail; modified gamma mgamma_range; power law its
// support [ln r_min, ln r_max]; monodisperse its radius (a range of width 0).
SCATRS_HD void radius_range(int kind, const double *p, double *t_lo, double *t_hi) {
    if (kind == kLogNormal) {
        const double c = log(p[0]), h = 5.0 * log(p[1]);
        *t_lo = c - h;
        *t_hi = c + h;
    } else if (kind == kGamma) {
        const double s = sqrt(p[1] / (1.0 - 2.0 * p[1]));
        const double c = log(p[0]);
        *t_lo = c - 8.0 * s;
        *t_hi = c + log(1.0 + kGammaTail * p[1] + sqrt(2.0 * kGammaTail * p[1]));
    } else if (kind == kModifiedGamma) {
        mgamma_range(p, t_lo, t_hi, nullptr, nullptr);
    } else if (kind == kPowerLaw) {
        *t_lo = log(p[1]);
        *t_hi = log(p[2]);
    } else if (kind == kSegment) {
        *t_lo = log(p[0]);
        *t_hi = log(p[1]);
    } else {
        *t_lo = *t_hi = log(p[0]);
    }
}

// The quadrature weight of a radius node: w (t_hi - t_lo)/2 u(r) r, and w r for a monodisperse
// law (every node at its radius, so that the averages are the sphere's).
SCATRS_HD double node_weight(int kind, double w, double half, double u, double r) {
    return w * (kind == kMonodisperse ? 1.0 : half) * u * r;
}

// Derivatives of the range's midpoint and half-width in ln r with respect to the parameters,
// and the half-width itself (a power law's limits are fixed: its range does not move).
SCATRS_HD double range_derivatives(int kind, const double *p, double dmid[3], double dhalf[3]) {
    for (int k = 0; k < 3; ++k) dmid[k] = dhalf[k] = 0.0;
    if (kind == kLogNormal) {
        dmid[0] = 1.0 / p[0];
        dhalf[1] = 5.0 / p[1];
        return 5.0 * log(p[1]);
    }
    if (kind == kGamma) {
        const double q = 1.0 - 2.0 * p[1];
        const double s = sqrt(p[1] / q);
        const double ds = 0.5 / (s * q * q);  // d s / d b
        const double rt = sqrt(2.0 * kGammaTail * p[1]);
        const double up = log(1.0 + kGammaTail * p[1] + rt);  // t_hi - ln a
        const double dup = (kGammaTail + kGammaTail / rt) / (1.0 + kGammaTail * p[1] + rt);
        dmid[0] = 1.0 / p[0];
        dmid[1] = 0.5 * (dup - 8.0 * ds);
        dhalf[1] = 0.5 * (dup + 8.0 * ds);
        return 0.5 * (up + 8.0 * s);
    }
    if (kind == kModifiedGamma) {
        double t_lo, t_hi, dlo[3], dhi[3];
        mgamma_range(p, &t_lo, &t_hi, dlo, dhi);
        for (int k = 0; k < 3; ++k) {
            dmid[k] = 0.5 * (dhi[k] + dlo[k]);
            dhalf[k] = 0.5 * (dhi[k] - dlo[k]);
        }
        return 0.5 * (t_hi - t_lo);
    }
    if (kind == kPowerLaw) return 0.5 * (log(p[2]) - log(p[1]));
    if (kind == kSegment) return 0.5 * (log(p[1]) - log(p[0]));
    dmid[0] = 1.0 / p[0];  // monodisperse: every node at the radius
    return 0.0;
}

// d(ln u)/d(ln r) at fixed parameters.
SCATRS_HD double log_density_slope(int kind, const double *p, double r) {
    if (kind == kLogNormal) {
        const double s = log(p[1]);
        return -(log(r) - log(p[0])) / (s * s) - 1.0;
    }
    if (kind == kGamma) return (1.0 - 3.0 * p[1]) / p[1] - r / (p[0] * p[1]);
    if (kind == kModifiedGamma) return p[0] - p[1] * p[2] * pow(r, p[2]);
    if (kind == kPowerLaw) return -p[0];
    if (kind == kSegment) {
        const double w = p[1] - p[0], s = (r - p[0]) / w;
        return r * (2.0 * p[2] - 1.0) / (w * ((1.0 - p[2]) * (1.0 - s) + p[2] * s));
    }
    return 0.0;
}

// ln u(r) and, with d not null, d(ln u)/dp[3], with the reference radius held fixed
// (distribution.rs:366-430); 0 for fixed and unused parameters.
SCATRS_HD double log_density(int kind, const double *p, double r, double *d) {
    const double r0 = dist_reference(kind, p);
    double ln_u = 0.0, d0 = 0.0, d1 = 0.0, d2 = 0.0;
    if (kind == kLogNormal) {
        const double mu = log(p[0]), s = log(p[1]);
        const double z = log(r) - mu;
        d0 = z / (s * s * p[0]);
        d1 = (z * z / (s * s * s) - 1.0 / s) / p[1];
        ln_u = -z * z / (2.0 * s * s) - log(s) - log(r / r0);
    } else if (kind == kGamma) {
        const double a = p[0], b = p[1];
        const double alpha = (1.0 - 3.0 * b) / b;
        const double scale = a * b;
        const double l = log(r / r0);
        d0 = (r - r0) * b / (scale * scale);
        d1 = -l / (b * b) + (r - r0) * a / (scale * scale);
        ln_u = alpha * l - (r - r0) / scale;
    } else if (kind == kModifiedGamma) {
        const double alpha = p[0], b = p[1], g = p[2];
        const double rg = pow(r, g), r0g = pow(r0, g);
        const double l = log(r / r0);
        d0 = l;
        d1 = -(rg - r0g);
        d2 = -b * (rg * log(r) - r0g * log(r0));
        ln_u = alpha * l - b * (rg - r0g);
    } else if (kind == kPowerLaw) {
        const double l = log(r / r0);
        d0 = -l;
        ln_u = -p[0] * l;
    } else if (kind == kSegment) {
        const double s = (r - p[0]) / (p[1] - p[0]);
        ln_u = log((1.0 - p[2]) * (1.0 - s) + p[2] * s);
    }
    if (d) {
        d[0] = d0;
        d[1] = d1;
        d[2] = d2;
    }
    return ln_u;
}

}  // namespace scatrs
