// Input checks of a sphere, as Mie::compute makes them (scatrs/src/mie/mod.rs:274-320) with the
// limits of mie/mod.rs:95-111 and the validated domain of mie/domain.rs:123-140: a device form
// that returns the code, and a host form that also writes the reference's message.
#pragma once

#include <cmath>
#include <cstdio>
#include <string>

#include "context.h"
#include "mie_device.h"

namespace scatrs {

constexpr double kMinSize = 1e-30, kMaxSize = 1e6;
constexpr double kMinIndex = 1e-6, kMaxIndex = 1e3;
constexpr long long kMaxOrder = 1ll << 20;
constexpr int64_t kMaxPhaseCosines = 1 << 16;  // scatrs_set_phase_cosines
constexpr double kDomXMin = 1e-6, kDomXMax = 2e4, kDomMReMax = 10.0, kDomMImMax = 10.0,
                 kDomAbsMMin = 0.1;

SCATRS_HD int sphere_code(double x, double mr, double mi, long long n_fixed, int allow,
                          long long *n_out) {
    *n_out = 0;
    if (!(x >= kMinSize && x <= kMaxSize)) return kInvalid;
    if (!(__builtin_isfinite(mr) && __builtin_isfinite(mi))) return kInvalid;
    if (mr < 0.0 || mi < 0.0) return kInvalid;
    const double abs_m = hypot(mr, mi);
    if (!(abs_m >= kMinIndex && abs_m <= kMaxIndex)) return kInvalid;
    const long long def = default_n_max(x);
    const long long n = n_fixed > 0 ? n_fixed : def;
    if (n == 0 || n > kMaxOrder) return kInvalid;
    if (!allow) {
        const bool in = x >= kDomXMin && x <= kDomXMax && mr > 0.0 && mr <= kDomMReMax &&
                        mi >= 0.0 && mi <= kDomMImMax && abs_m >= kDomAbsMMin;
        if (!in || n < def) return kOutside;
    }
    *n_out = n;
    return kValid;
}

inline std::string fmt_e(double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%e", v);
    return b;
}

// The code and the reference's message for one sphere (host).
inline int sphere_message(double x, double mr, double mi, long long n_fixed, int allow,
                          std::string *msg) {
    if (!(x >= kMinSize && x <= kMaxSize)) {
        *msg = "the size parameter x must be in [1e-30, 1e6], got " + fmt_e(x);
        return kInvalid;
    }
    if (!(std::isfinite(mr) && std::isfinite(mi))) {
        *msg = "the refractive index must be finite, got " + fmt_e(mr) + " + " + fmt_e(mi) + "i";
        return kInvalid;
    }
    if (mr < 0.0 || mi < 0.0) {
        *msg = "the refractive index must have Re m >= 0 and Im m >= 0 (Im m < 0 is a gain "
               "medium), got " + fmt_e(mr) + " + " + fmt_e(mi) + "i";
        return kInvalid;
    }
    const double abs_m = std::hypot(mr, mi);
    if (!(abs_m >= kMinIndex && abs_m <= kMaxIndex)) {
        *msg = "|m| must be in [1e-6, 1e3], got " + fmt_e(abs_m);
        return kInvalid;
    }
    const long long def = default_n_max(x);
    const long long n = n_fixed > 0 ? n_fixed : def;
    if (n == 0 || n > kMaxOrder) {
        *msg = "the number of terms must be in [1, 1048576], got " + std::to_string(n);
        return kInvalid;
    }
    if (!allow) {
        const bool in = x >= kDomXMin && x <= kDomXMax && mr > 0.0 && mr <= kDomMReMax &&
                        mi >= 0.0 && mi <= kDomMImMax && abs_m >= kDomAbsMMin;
        if (!in) {
            *msg = "x = " + fmt_e(x) + ", m = " + fmt_e(mr) + " + " + fmt_e(mi) +
                   "i is outside the validated domain x in [1e-6, 2e4], 0 < Re m <= 10, "
                   "0 <= Im m <= 10, |m| >= 0.1";
            return kOutside;
        }
        if (n < def) {
            *msg = "n_max = " + std::to_string(n) + " is below the default truncation " +
                   std::to_string(def) + " for x = " + fmt_e(x);
            return kOutside;
        }
    }
    *msg = "";
    return kValid;
}


// Checks of one layered sphere given its interface size parameters x[0..L-1] (outer last) and
// indices m (interleaved re, im per layer), as Layered::compute / compute_with_derivatives make
// them (scatrs/src/layered/mod.rs:159-190,229-319) with the validated domain of
// layered/domain.rs:53-95: x_L in [1e-6, 1e4], every layer 0 < Re m <= 10, 0 <= Im m <= 10,
// |m| >= 0.1 and |m| x_L <= 3e4, at most 16 layers; with derivatives x_L <= 1e3. One layer is
// the sphere, with the sphere's checks.
constexpr int kDomMaxLayers = 16;
constexpr double kDomLayeredXMax = 1e4, kDomLayeredPhase = 3e4, kDomLayeredDerivXMax = 1e3;

SCATRS_HD int layered_x_code(const double *x, const double *m, int L, long long n_fixed,
                             int allow, int derivs, long long *n_out) {
    *n_out = 0;
    if (L == 1) return sphere_code(x[0], m[0], m[1], n_fixed, allow, n_out);
    for (int l = 0; l < L; ++l)
        if (!__builtin_isfinite(x[l])) return kInvalid;
    for (int l = 1; l < L; ++l)
        if (x[l] < x[l - 1]) return kInvalid;
    const double outer = x[L - 1];
    if (x[0] < kMinSize || outer > kMaxSize) return kInvalid;
    for (int l = 0; l < L; ++l) {
        const double mr = m[2 * l], mi = m[2 * l + 1];
        if (!(__builtin_isfinite(mr) && __builtin_isfinite(mi))) return kInvalid;
        if (mr < 0.0 || mi < 0.0) return kInvalid;
        const double am = hypot(mr, mi);
        if (!(am >= kMinIndex && am <= kMaxIndex)) return kInvalid;
    }
    const long long def = default_n_max(outer);
    const long long n = n_fixed > 0 ? n_fixed : def;
    if (n == 0 || n > kMaxOrder) return kInvalid;
    if (!allow) {
        bool in = L <= kDomMaxLayers && outer >= kDomXMin && outer <= kDomLayeredXMax;
        for (int l = 0; l < L; ++l) {
            const double mr = m[2 * l], mi = m[2 * l + 1], am = hypot(mr, mi);
            in = in && am * outer <= kDomLayeredPhase && mr > 0.0 && mr <= kDomMReMax &&
                 mi >= 0.0 && mi <= kDomMImMax && am >= kDomAbsMMin;
        }
        if (!in || n < def) return kOutside;
        if (derivs && outer > kDomLayeredDerivXMax) return kOutside;
    }
    *n_out = n;
    return kValid;
}

// The code and the reference's message for one layered sphere (host).
inline int layered_x_message(const double *x, const double *m, int L, long long n_fixed,
                             int allow, int derivs, std::string *msg) {
    if (L == 1) return sphere_message(x[0], m[0], m[1], n_fixed, allow, msg);
    for (int l = 0; l < L; ++l)
        if (!std::isfinite(x[l])) {
            *msg = "the size parameter of layer " + std::to_string(l) + " must be finite, got " +
                   fmt_e(x[l]);
            return kInvalid;
        }
    for (int l = 1; l < L; ++l)
        if (x[l] < x[l - 1]) {
            *msg = "the size parameters must be nondecreasing from the core out, but layer " +
                   std::to_string(l) + " has " + fmt_e(x[l]) + " < " + fmt_e(x[l - 1]);
            return kInvalid;
        }
    const double outer = x[L - 1];
    if (x[0] < kMinSize || outer > kMaxSize) {
        *msg = "the size parameters must be in [1e-30, 1e6], got " + fmt_e(x[0]) + " to " +
               fmt_e(outer);
        return kInvalid;
    }
    std::string ms = "[";
    for (int l = 0; l < L; ++l) {
        const double mr = m[2 * l], mi = m[2 * l + 1];
        const std::string lay = "layer " + std::to_string(l);
        if (!(std::isfinite(mr) && std::isfinite(mi))) {
            *msg = "the refractive index of " + lay + " must be finite, got " + fmt_e(mr) +
                   " + " + fmt_e(mi) + "i";
            return kInvalid;
        }
        if (mr < 0.0 || mi < 0.0) {
            *msg = "the refractive index of " + lay +
                   " must have Re m >= 0 and Im m >= 0 (Im m < 0 is a gain medium), got " +
                   fmt_e(mr) + " + " + fmt_e(mi) + "i";
            return kInvalid;
        }
        const double am = std::hypot(mr, mi);
        if (!(am >= kMinIndex && am <= kMaxIndex)) {
            *msg = "|m| of " + lay + " must be in [1e-6, 1e3], got " + fmt_e(am);
            return kInvalid;
        }
        ms += (l ? ", " : "") + fmt_e(mr) + " + " + fmt_e(mi) + "i";
    }
    ms += "]";
    const long long def = default_n_max(outer);
    const long long n = n_fixed > 0 ? n_fixed : def;
    if (n == 0 || n > kMaxOrder) {
        *msg = "the number of terms must be in [1, 1048576], got " + std::to_string(n);
        return kInvalid;
    }
    if (!allow) {
        bool in = L <= kDomMaxLayers && outer >= kDomXMin && outer <= kDomLayeredXMax;
        for (int l = 0; l < L; ++l) {
            const double mr = m[2 * l], mi = m[2 * l + 1], am = std::hypot(mr, mi);
            in = in && am * outer <= kDomLayeredPhase && mr > 0.0 && mr <= kDomMReMax &&
                 mi >= 0.0 && mi <= kDomMImMax && am >= kDomAbsMMin;
        }
        if (!in) {
            *msg = "a layered sphere of " + std::to_string(L) + " layers with x_L = " +
                   fmt_e(outer) + " and m = " + ms +
                   " is outside the validated domain x_L in [1e-6, 1e4], every layer 0 < Re m "
                   "<= 10, 0 <= Im m <= 10, |m| >= 0.1 and |m| x_L <= 3e4, at most 16 layers";
            return kOutside;
        }
        if (n < def) {
            *msg = "n_max = " + std::to_string(n) + " is below the default truncation " +
                   std::to_string(def) + " for x_L = " + fmt_e(outer);
            return kOutside;
        }
        if (derivs && outer > kDomLayeredDerivXMax) {
            *msg = "the derivatives are validated up to x_L = 1e3, got " + fmt_e(outer);
            return kOutside;
        }
    }
    *msg = "";
    return kValid;
}

}  // namespace scatrs
