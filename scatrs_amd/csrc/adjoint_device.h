// The per-node coefficient adjoint of the ensemble path's reverse mode: a sink for the
// Mie / layered order loop (mie_orders) that contracts each order's coefficient cotangents with
// the coefficient derivatives it receives.
#pragma once

#include "layered_device.h"
#include "mie_device.h"

namespace scatrs {

// Contracts the cotangents of a node's a_n, b_n with their derivatives, order by order, for
// the outer layer (order_d: d/dx, d/dm of the outer layer) and, for a layered sphere, every
// inner layer (stored: from the inner-layer sensitivities of the order, as
// layered/kernel.rs:549-587 forms them). The cotangent of a_n is
//   tfac abar_T + 2 alpha (2n+1) a_n + beta (2n+1) (-1)^n B
//       + gamma (pair(n-1) a_{n-1} + pair(n) a_{n+1} + own(n) b_n)
// (b_n: the same with a, b swapped and -beta), from C_sca = (lambda^2/2pi) sum (2n+1)(|a|^2+|b|^2),
// C_back = (lambda^2/4pi) |B|^2 with B = sum (2n+1)(-1)^n (a_n - b_n) and
// C_sca g = (lambda^2/pi) G, G = sum pair(n-1) Re(a_{n-1} conj a_n + ...) + own(n) Re(a_n conj b_n)
// (efficiencies.rs:64-84), pair(k) = k (k+2)/(k+1), own(n) = (2n+1)/(n(n+1)). a_{n+1}, b_{n+1}
// are read from the forward pass's staging buffer.
template <int L>
struct AdjointSink {
    const cx *sa, *sb;  // staged a_n, b_n (index 2 (n - 1): interleaved)
    const cx *ta, *tb;  // table cotangents of a_n, b_n (tfac = 0 and any finite n values: none)
    int n;
    double tfac, alpha, beta, gamma;
    cx back;
    cx a_prev, b_prev, cur_a, cur_b;
    double tx[L];
    cx tm[L];
    double jabs[L][3];
    double abs_sum;

    __device__ void init() {
        a_prev = b_prev = cur_a = cur_b = real(0.0);
#pragma unroll
        for (int l = 0; l < L; ++l) {
            tx[l] = 0.0;
            tm[l] = real(0.0);
            jabs[l][0] = jabs[l][1] = jabs[l][2] = 0.0;
        }
        abs_sum = 0.0;
    }

    __device__ void order(int, cx, cx) {}

    __device__ void order_d(int k, cx a, cx b, cx dax, cx dam, cx dbx, cx dbm) {
        const double kf = (double)k, w = (double)(2 * k + 1);
        const cx a1 = k < n ? sa[2 * k] : real(0.0), b1 = k < n ? sb[2 * k] : real(0.0);
        const double pp = k >= 2 ? (kf - 1.0) * (kf + 1.0) / kf : 0.0;
        const double pn = k < n ? kf * (kf + 2.0) / (kf + 1.0) : 0.0;
        const double own = w / (kf * (kf + 1.0));
        const double bs = (k & 1) ? -(beta * w) : beta * w;
        const cx ca = a * (2.0 * alpha * w) + back * bs + (a_prev * pp + a1 * pn + b * own) * gamma +
                      ta[k - 1] * tfac;
        const cx cb = b * (2.0 * alpha * w) - back * bs + (b_prev * pp + b1 * pn + a * own) * gamma +
                      tb[k - 1] * tfac;
        tx[L - 1] += re_mul_conj(ca, dax) + re_mul_conj(cb, dbx);
        tm[L - 1] = tm[L - 1] + conj(ca) * dam + conj(cb) * dbm;
        a_prev = a;
        b_prev = b;
        cur_a = ca;
        cur_b = cb;
    }

    template <class O>
    __device__ void stored(int k, cx inv_da, cx inv_db, cx g_a, cx g_b, cx d_m, cx md,
                           const O &o) {
        if constexpr (L > 1) {
            const cx mL = o.m[L - 1];
            const cx inv_mL = inv(mL);
            const cx ida2 = inv_da * inv_da, idb2 = inv_db * inv_db;
            const double na = norm2(inv_da), nb = norm2(inv_db);
            const double w = (double)(2 * k + 1);
#pragma unroll
            for (int l = 0; l < L - 1; ++l) {
                const cx ba_x = o.dh[l][0] * inv_mL, ba_m = o.dh[l][1] * inv_mL;
                const cx bb_x = mL * o.dh[l][2], bb_m = mL * o.dh[l][3];
                const cx dax = -(times_i(ba_x) * ida2), dam = -(times_i(ba_m) * ida2);
                const cx dbx = -(times_i(bb_x) * idb2), dbm = -(times_i(bb_m) * idb2);
                tx[l] += re_mul_conj(cur_a, dax) + re_mul_conj(cur_b, dbx);
                tm[l] = tm[l] + conj(cur_a) * dam + conj(cur_b) * dbm;
                jabs[l][0] += w * (absorption_term(d_m, ba_x, ba_x * g_a, na) +
                                   absorption_term(md, bb_x, bb_x * g_b, nb));
                jabs[l][1] += w * (absorption_term(d_m, ba_m, ba_m * g_a, na) +
                                   absorption_term(md, bb_m, bb_m * g_b, nb));
                const cx iba = times_i(ba_m), ibb = times_i(bb_m);
                jabs[l][2] += w * (absorption_term(d_m, iba, iba * g_a, na) +
                                   absorption_term(md, ibb, ibb * g_b, nb));
            }
        } else {
            (void)k, (void)inv_da, (void)inv_db, (void)g_a, (void)g_b, (void)d_m, (void)md, (void)o;
        }
    }

    __device__ void done(double s, double j0, double j1, double j2) {
        abs_sum = s;
        jabs[L - 1][0] = j0;
        jabs[L - 1][1] = j1;
        jabs[L - 1][2] = j2;
    }

    // A lossless layered sphere stays lossless as any x_l or Re m_l changes
    // (layered/kernel.rs:453-464).
    __device__ void zero_absorption() {
        abs_sum = 0.0;
#pragma unroll
        for (int l = 0; l < L; ++l) jabs[l][0] = jabs[l][1] = 0.0;
    }
};

}  // namespace scatrs
