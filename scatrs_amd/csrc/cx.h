// Complex FP64 arithmetic for the device recurrences: a plain (re, im) pair with the
// operations the Mie and layered recurrences need, written out so the compiler keeps each
// product as two FMAs-free products and one add (the recurrences' accuracy model assumes one
// rounding per operation, like the reference's num-complex arithmetic).
#pragma once

#include <hip/hip_runtime.h>

namespace scatrs {

struct cx {
    double re, im;
};

#define SCATRS_HD __host__ __device__ __forceinline__

SCATRS_HD cx mk(double re, double im) { return cx{re, im}; }
SCATRS_HD cx real(double re) { return cx{re, 0.0}; }
SCATRS_HD cx operator+(cx a, cx b) { return cx{a.re + b.re, a.im + b.im}; }
SCATRS_HD cx operator-(cx a, cx b) { return cx{a.re - b.re, a.im - b.im}; }
SCATRS_HD cx operator-(cx a) { return cx{-a.re, -a.im}; }
SCATRS_HD cx operator*(cx a, cx b) {
    return cx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
SCATRS_HD cx operator*(cx a, double t) { return cx{a.re * t, a.im * t}; }
SCATRS_HD cx operator*(double t, cx a) { return cx{a.re * t, a.im * t}; }
SCATRS_HD cx operator+(cx a, double t) { return cx{a.re + t, a.im}; }
SCATRS_HD cx operator+(double t, cx a) { return cx{t + a.re, a.im}; }
SCATRS_HD cx operator-(cx a, double t) { return cx{a.re - t, a.im}; }
SCATRS_HD cx operator-(double t, cx a) { return cx{t - a.re, -a.im}; }
SCATRS_HD cx times_i(cx u) { return cx{-u.im, u.re}; }
SCATRS_HD cx conj(cx a) { return cx{a.re, -a.im}; }
SCATRS_HD double norm2(cx a) { return a.re * a.re + a.im * a.im; }
SCATRS_HD double norm1(cx a) { return fabs(a.re) + fabs(a.im); }
SCATRS_HD double cabs(cx a) { return hypot(a.re, a.im); }
SCATRS_HD bool is_zero(cx a) { return a.re == 0.0 && a.im == 0.0; }
SCATRS_HD double re_mul_conj(cx u, cx v) { return u.re * v.re + u.im * v.im; }

// 1/s for s of normal magnitude: on the device v_rcp_f64 refined by two Newton steps
// (r <- r + r (1 - s r)), within an ulp of the quotient and about half the instructions of the
// IEEE division sequence; on the host the plain quotient.
SCATRS_HD double recip(double s) {
#if defined(__HIP_DEVICE_COMPILE__)
    double r = __builtin_amdgcn_rcp(s);
    double e = fma(-s, r, 1.0);
    r = fma(r, e, r);
    e = fma(-s, r, 1.0);
    return fma(r, e, r);
#else
    return 1.0 / s;
#endif
}

// 1/z with one division inside 1e-300 < |z|^2 < 1e300 (d != 0), Smith's two-division form
// outside that range or for real z, where it is exactly 1/c (scatrs/src/complex.rs:21-38).
SCATRS_HD cx inv(cx z) {
    const double c = z.re, d = z.im;
    const double s = c * c + d * d;
    if (s > 1e-300 && s < 1e300 && d != 0.0) {
        const double t = recip(s);
        return cx{c * t, -(d * t)};
    }
    if (fabs(d) <= fabs(c)) {
        const double r = d / c;
        const double t = 1.0 / (c + d * r);
        return cx{t, -(r * t)};
    }
    const double r = c / d;
    const double t = 1.0 / (c * r + d);
    return cx{r * t, -t};
}

// e^w - 1 for Re w <= 0 (scatrs/src/complex.rs:51-56).
SCATRS_HD cx exp_m1(cx w) {
    double s, c;
    sincos(w.im, &s, &c);
    const double h = sin(w.im * 0.5);
    return cx{expm1(w.re) * c - h * h * 2.0, exp(w.re) * s};
}

SCATRS_HD cx csin(cx z) { return cx{sin(z.re) * cosh(z.im), cos(z.re) * sinh(z.im)}; }
SCATRS_HD cx ccos(cx z) { return cx{cos(z.re) * cosh(z.im), -(sin(z.re) * sinh(z.im))}; }

}  // namespace scatrs
