// Host-side quadrature tables of the ensemble path: Gauss-Legendre rules and the Wigner d
// functions of the Greek projection. Computed once per size on the host in extended precision
// and uploaded; the device kernels only read them.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <thread>
#include <vector>

namespace scatrs {

// n-point Gauss-Legendre rule on [-1, 1], nodes ascending and exactly symmetric
// (scatrs/src/special/gauss_legendre.rs:210-216 gives the same rule): Newton's method on the
// three-term recurrence of P_n in long double from Tricomi's initial guesses. Each node costs
// O(n); rules of 512 points or more are computed on up to 16 host threads (started and joined
// here), each taking every t-th node, with the same arithmetic per node whatever the count.
inline void gauss_legendre(int64_t n, std::vector<double> &nodes, std::vector<double> &weights) {
    nodes.assign(n, 0.0);
    weights.assign(n, 0.0);
    if (n == 1) {
        weights[0] = 2.0;
        return;
    }
    const long double pi = 3.14159265358979323846264338327950288L;
    auto legendre = [n](long double x, long double &p, long double &dp) {
        long double p0 = 1.0L, p1 = x;
        for (int64_t k = 2; k <= n; ++k) {
            const long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / (long double)k;
            p0 = p1;
            p1 = p2;
        }
        p = p1;
        dp = (long double)n * (x * p1 - p0) / (x * x - 1.0L);
    };
    auto node = [&](int64_t i) {
        long double x = cosl(pi * ((long double)i + 0.75L) / ((long double)n + 0.5L));
        long double p = 0, dp = 1;
        for (int it = 0; it < 100; ++it) {
            legendre(x, p, dp);
            const long double dx = p / dp;
            x -= dx;
            if (fabsl(dx) <= 1e-19L * fabsl(x) || fabsl(dx) < 1e-21L) break;
        }
        if (n % 2 == 1 && i == (n - 1) / 2) x = 0.0L;
        legendre(x, p, dp);
        const long double w = 2.0L / ((1.0L - x * x) * dp * dp);
        nodes[n - 1 - i] = (double)x;
        nodes[i] = -(double)x;
        weights[n - 1 - i] = (double)w;
        weights[i] = (double)w;
    };
    const int64_t half = (n + 1) / 2;
    const int64_t threads =
        n < 512 ? 1 : std::min<int64_t>({16, half, std::max(1u, std::thread::hardware_concurrency())});
    if (threads <= 1) {
        for (int64_t i = 0; i < half; ++i) node(i);
        return;
    }
    std::vector<std::thread> pool;
    for (int64_t t = 0; t < threads; ++t)
        pool.emplace_back([&node, t, threads, half] {
            for (int64_t i = t; i < half; i += threads) node(i);
        });
    for (auto &th : pool) th.join();
}

// d^l_{mn}(x) for l = 0..lmax into out (scatrs/src/special/wigner.rs:152-182), for the small
// |m|, |n| of the projection ((0,0), (2,2), (2,-2), (0,2)): the upward three-term recurrence
// (Mishchenko, Travis & Lacis 2002, Eq. B.22) away from the poles, Reinsch's difference form
// on d^s - d^{s-1} next to them (|x| >= 0.9, or m = +-n on the side where |d| -> 1), with
// d^l_{mn}(x) = (-1)^{l+m} d^l_{m,-n}(-x) mapping x < 0 onto x > 0.
inline void wigner_d_small(int m, int n, double x, int64_t lmax, double *out) {
    const int64_t len = lmax + 1;
    const int64_t l_min = std::max(std::abs(m), std::abs(n));
    for (int64_t l = 0; l < std::min(l_min, len); ++l) out[l] = 0.0;
    if (l_min >= len) return;
    const bool reflect = x < 0.0 && (m == -n || -x >= 0.9);
    if (reflect) {
        n = -n;
        x = -x;
    }
    const int64_t a = std::abs(m - n), b = std::abs(m + n);
    // sqrt(C(a + b, a)) sin^a(theta/2) cos^b(theta/2)
    const double hs = std::sqrt((1.0 - x) * 0.5), hc = std::sqrt((1.0 + x) * 0.5);
    double start = 1.0;
    for (int64_t k = 1; k <= std::max(a, b); ++k) {
        if (k <= a) start = start * std::sqrt((double)(b + k) / (double)k) * hs;
        if (k <= b) start *= hc;
    }
    if (!(n >= m || (m - n) % 2 == 0)) start = -start;
    const double mf = m, nf = n;
    if (m == n && x >= 0.0) {
        const double w = x - 1.0;
        int64_t s;
        double delta, cur;
        if (l_min == 0) {
            out[0] = start;
            if (len == 1) goto done;
            s = 1;
            delta = w;
            cur = x;
        } else {
            s = l_min;
            delta = start;
            cur = start;
        }
        out[s] = cur;
        while (s + 1 < len) {
            const double sf = (double)s, s1 = sf + 1.0;
            const double k = (2.0 * sf + 1.0) * sf * s1;
            const double bb = s1 * ((sf - mf) * (sf + mf));
            const double c = sf * ((s1 - mf) * (s1 + mf));
            const double next = (w * k * cur + delta * bb) * (1.0 / c);
            delta = next;
            cur += next;
            out[++s] = cur;
        }
    } else if (x >= 0.9) {
        const double w = x - 1.0, lf = (double)l_min;
        double q = 0.0, u = lf * (lf + 1.0) - mf * nf, delta = start, cur = start;
        int64_t s = l_min;
        out[s] = cur;
        while (s + 1 < len) {
            const double sf = (double)s, s1 = sf + 1.0;
            const double q1 = std::sqrt(((s1 - mf) * (s1 + mf)) * ((s1 - nf) * (s1 + nf)));
            const double alpha = sf * s1 - mf * nf;
            const double dmn = mf - nf;
            const double numer = (s1 * s1) * (1.0 + dmn * dmn - 2.0 * s1) + (2.0 * s1) * (mf * nf);
            const double v = numer / (alpha + q1);
            const double k = w * ((2.0 * sf + 1.0) * sf * s1) + u * s1 + v * sf;
            const double next = (k * cur + delta * (q * s1)) * (1.0 / (q1 * sf));
            q = q1;
            u = v + 2.0 * s1;
            delta = next;
            cur += next;
            out[++s] = cur;
        }
    } else {
        double q = 0.0, prev = 0.0, cur = start;
        int64_t s = l_min;
        out[s] = cur;
        while (s + 1 < len) {
            const double sf = (double)s;
            const double c_prev = q * (sf + 1.0);
            const double numer = (x * (sf * (sf + 1.0)) - mf * nf) * (2.0 * sf + 1.0) * cur - c_prev * prev;
            const double s1 = sf + 1.0;
            q = std::sqrt((s1 - mf) * (s1 + mf)) * std::sqrt((s1 - nf) * (s1 + nf));
            prev = cur;
            cur = numer / (q * sf);
            out[++s] = cur;
        }
    }
done:
    if (reflect) {
        int64_t r = (l_min + m) % 2;
        if (r < 0) r += 2;
        for (int64_t i = 1 - r; l_min + i < len; i += 2) out[l_min + i] = -out[l_min + i];
    }
}

}  // namespace scatrs
