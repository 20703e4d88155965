// Per-lane Lorenz-Mie coefficient solver for gfx950: one sphere per lane of a 64-wide
// wavefront. The downward ratio chains u_k(mx) and u_k(x) are written to a lane-strided
// scratch column (element k of lane l at base[k * stride + l], so a wavefront's 64 lanes touch
// 64 consecutive doubles per order) and read back by the upward order loop, which keeps
// psi, chi, T_n, 1/xi_n, R_n and the running sums in registers.
//
// The algorithm is the reference's (scatrs/src/mie/kernel.rs:99-559): the same start index,
// the same ratio-region numerators free of cancellation, the same absorption identity and the
// same closed-form derivatives; only the summation order of the efficiency series differs
// (the g pair term is added when order k+1 arrives).
#pragma once

#include <type_traits>

#include "cx.h"

namespace scatrs {

constexpr double kEps = 2.220446049250313080847e-16;  // 2^-52
constexpr double kBits = 52.0;
constexpr int kMargin = 30;  // ceil(bits / 2) + 4 (mie/kernel.rs:82-84)


// N = ceil(x + 6 x^(1/3) + 3) (scatrs/src/mie/mod.rs:122-124).
SCATRS_HD long long default_n_max(double x) {
    const double n = ceil(x + 6.0 * cbrt(x) + 3.0);
    return (n < 9.0e15) ? (long long)n : (long long)9.0e15;
}

// The first index of a downward ratio recurrence reaching n_top with the start value
// forgotten to rounding level (scatrs/src/mie/kernel.rs:99-112).
SCATRS_HD long long start_index(long long n_top, double rho_abs) {
    long long k0 = n_top;
    const long long c15 = (long long)ceil(1.5 * rho_abs);
    if (c15 > k0) k0 = c15;
    const double q = (double)(2 * k0 + 1) / rho_abs - 0.5;
    const long long enclosed = k0 + (long long)ceil(kBits / (2.0 * log2(q))) + 2;
    if ((double)n_top >= 1.5 * rho_abs) return enclosed;
    const double s0 = (double)n_top - rho_abs;
    const double s0_32 = s0 > 0.0 ? s0 * sqrt(s0) : 0.0;
    const double s = pow(s0_32 + 1.4 * 0.37 * kBits * sqrt(rho_abs), 2.0 / 3.0);
    const long long asymptotic = (long long)ceil(rho_abs + s) + kMargin;
    return asymptotic < enclosed ? asymptotic : enclosed;
}

SCATRS_HD int upward_orders(double x, int n) {
    if (x >= 1.0) {
        const double f = floor(x);
        return f < (double)n ? (int)f : n;
    }
    return 0;
}

__device__ __forceinline__ cx step_complex(int k, cx inv_z, cx u) {
    const cx c = inv_z * (double)(2 * k + 1);
    cx den = c - u;
    if (is_zero(den)) den = real(kEps * (cabs(c) + cabs(u)));
    return inv(den);
}

__device__ __forceinline__ double step_real(int k, double inv_rho, double u) {
    const double c = inv_rho * (double)(2 * k + 1);
    double den = c - u;
    if (den == 0.0) den = kEps * (fabs(c) + fabs(u));
    return recip(den);
}

// The plain-path check of a block of complex reciprocals: the smallest and largest squared
// modulus and the smallest |Im| of the arguments. inv takes its one-division form exactly when
// 1e-300 < |z|^2 < 1e300 and Im z != 0 (and step_complex keeps its denominator when it is not
// 0), so a block whose arguments all pass computes, without any branch, what inv would.
struct PlainCheck {
    double lo = 1e308, hi = 0.0, im = 1e308;
    __device__ __forceinline__ bool ok() const { return lo > 1e-300 && hi < 1e300 && im != 0.0; }
    __device__ __forceinline__ bool ok_pair() const { return lo > 1e-150 && hi < 1e150; }
};

// 1/z by the one-division form of inv, its argument noted in pc.
__device__ __forceinline__ cx inv_plain(cx z, PlainCheck &pc) {
    const double c = z.re, d = z.im;
    const double s = c * c + d * d;
    pc.lo = fmin(pc.lo, s);
    pc.hi = fmax(pc.hi, s);
    pc.im = fmin(pc.im, fabs(d));
    const double t = recip(s);
    return cx{c * t, -(d * t)};
}

// 1/za and 1/zb from one reciprocal, 1/(|za|^2 |zb|^2), for two independent inverses of the
// same order: within a few ulps of inv's wherever both squared moduli lie in (1e-150, 1e150),
// which the block checks afterwards from pc (pc.ok_pair()).
// na, nb receive 1/|za|^2 and 1/|zb|^2.
__device__ __forceinline__ void inv_plain_pair(cx za, cx zb, PlainCheck &pc, cx &ia, cx &ib,
                                               double &na, double &nb) {
    const double sa = za.re * za.re + za.im * za.im;
    const double sb = zb.re * zb.re + zb.im * zb.im;
    pc.lo = fmin(pc.lo, fmin(sa, sb));
    pc.hi = fmax(pc.hi, fmax(sa, sb));
    const double t = recip(sa * sb);
    const double ta = sb * t, tb = sa * t;
    ia = cx{za.re * ta, -(za.im * ta)};
    ib = cx{zb.re * tb, -(zb.im * tb)};
    na = ta;
    nb = tb;
}

// step_complex on the plain path (its result is step_complex's when pc.ok() after the block).
__device__ __forceinline__ cx step_complex_plain(int k, cx inv_z, cx u, PlainCheck &pc) {
    const cx c = inv_z * (double)(2 * k + 1);
    return inv_plain(c - u, pc);
}

// u_k(m x) for k = 1..n_top+1 into col[k * stride] (scatrs/src/mie/kernel.rs:159-205).
__device__ __forceinline__ void ratio_chain(cx m, double x, int n_top, cx *col, int stride) {
    const cx z = m * x;
    if (m.im == 0.0) {
        const double inv_rho = 1.0 / z.re;
        const long long start = start_index(n_top + 1, fabs(z.re));
        double u = 0.0;
        for (long long k = start; k >= 1; --k) {
            u = step_real((int)k, inv_rho, u);
            if (k <= n_top + 1) col[k * stride] = real(u);
        }
    } else {
        const cx inv_z = inv(z);
        const long long start = start_index(n_top + 1, cabs(z));
        cx u = real(0.0);
        for (long long k = start; k >= 1; --k) {
            u = step_complex((int)k, inv_z, u);
            if (k <= n_top + 1) col[k * stride] = u;
        }
    }
}

// u_k(x) above the upward orders, k = max(n_s + 1, 1)..n+1 (scatrs/src/mie/kernel.rs:249-270).
__device__ __forceinline__ void x_chain(double x, int n, int n_s, double *col, int stride) {
    if (n_s >= n) return;
    const int top = n + 1;
    const long long start = start_index(top, fabs(x));
    const int stop = n_s + 1 > 1 ? n_s + 1 : 1;
    const double inv_x = 1.0 / x;
    double u = 0.0;
    for (long long k = start; k >= stop; --k) {
        u = step_real((int)k, inv_x, u);
        if (k <= top) col[k * stride] = u;
    }
}

// The order loop reads u_{k+1}(mx) in ascending k, after a downward pass from the start
// index. Two sources:
//  * ZColumn: the downward pass stored every value in a lane-strided global column;
//  * ZCheckpoint: the downward pass kept one value per block of kZBlock indices (the value
//    just above each block) in a short global column; the order loop regenerates each block
//    into the wavefront's LDS tile by the same downward steps before using it, so the full
//    column never goes through HBM.
constexpr int kZBlock = 16;
// Orders per checkpoint block of the single-particle batches (the sphere batch and the layered
// entry points' one-layer case share it, so that their results are the same bit for bit).
constexpr int kSphereBlock = 8;

struct ZColumn {
    static constexpr bool kBlocked = false;
    static constexpr bool kPlain = false;
    const cx *uz;
    int stride;
};

// PLAIN = true: every full block (of the downward pass, of the regeneration and of the orders
// below floor(x)) runs on the plain path without a redo, its arguments noted in pc (the
// downward steps) and pc_pair (the orders' paired inverses); the caller redoes the sphere on the
// checked path (PLAIN = false) in the rare lane whose checks fail.
template <int B, bool PLAIN = false>
struct ZCheckpointT {
    static constexpr bool kBlocked = true;
    static constexpr bool kPlain = PLAIN;
    static constexpr int kBlock = B;
    const cx *ckpt;     // ckpt[j * stride]: u at index 2 + (j + 1) B
    int stride;
    cx *lds;            // the lane's tile: element i at lds[i * 64]
    cx inv_z;           // 1/(mx), complex chain
    double inv_rho;     // 1/Re(mx), real chain
    bool real_chain;
    long long start;
    mutable PlainCheck pc, pc_pair;

    // Whether every plain step of the sphere stayed in the plain range (PLAIN only).
    __device__ __forceinline__ bool plain_ok() const { return pc.ok() && pc_pair.ok_pair(); }

    // u_i for i in [2 + j B, 2 + j B + B - 1] into lds[(i - 2 - j B) * 64], regenerated
    // downward from the checkpoint above the block (or from 0 above the start index).
    __device__ __forceinline__ void regen(int j) const {
        const long long lo = 2 + (long long)j * B;
        const long long top = lo + B - 1;
        const long long hi = top < start ? top : start;
        if (real_chain) {
            double u = (top + 1 <= start) ? ckpt[(int64_t)j * stride].re : 0.0;
            for (long long i = hi; i >= lo; --i) {
                u = step_real((int)i, inv_rho, u);
                lds[(i - lo) * 64] = real(u);
            }
        } else if (top + 1 <= start) {
            // A full block below the start: B steps without branches, redone on the checked
            // path in the rare lane whose arguments left the plain range.
            const cx u0 = ckpt[(int64_t)j * stride];
            cx u = u0;
            if constexpr (PLAIN) {
#pragma unroll
                for (int q = B - 1; q >= 0; --q) {
                    u = step_complex_plain((int)lo + q, inv_z, u, pc);
                    lds[q * 64] = u;
                }
                return;
            }
            PlainCheck pc;
#pragma unroll
            for (int q = B - 1; q >= 0; --q) {
                u = step_complex_plain((int)lo + q, inv_z, u, pc);
                lds[q * 64] = u;
            }
            if (!pc.ok()) {
                u = u0;
                for (int q = B - 1; q >= 0; --q) {
                    u = step_complex((int)lo + q, inv_z, u);
                    lds[q * 64] = u;
                }
            }
        } else {
            cx u = real(0.0);
            for (long long i = hi; i >= lo; --i) {
                u = step_complex((int)i, inv_z, u);
                lds[(i - lo) * 64] = u;
            }
        }
    }
};

using ZCheckpoint = ZCheckpointT<kZBlock>;

// The downward pass of u_k(m x) from the start index, keeping only the checkpoints of
// ZCheckpointT<B> (indices 2 + (j + 1) B for the blocks that hold orders 1..n).
template <int B = kZBlock, bool PLAIN = false>
__device__ __forceinline__ ZCheckpointT<B, PLAIN> ratio_chain_ckpt(cx m, double x, int n,
                                                                  cx *ckpt, int stride, cx *lds) {
    ZCheckpointT<B, PLAIN> zs;
    const cx z = m * x;
    zs.ckpt = ckpt;
    zs.stride = stride;
    zs.lds = lds;
    zs.real_chain = m.im == 0.0;
    const int jmax = (n - 1) / B;  // block of the last index used, n + 1
    const long long cmax = 2 + (long long)(jmax + 1) * B;
    const long long c0 = 2 + B;
    if (zs.real_chain) {
        zs.inv_rho = 1.0 / z.re;
        zs.inv_z = real(0.0);
        zs.start = start_index(n + 1, fabs(z.re));
        double u = 0.0;
        for (long long k = zs.start; k >= c0; --k) {
            u = step_real((int)k, zs.inv_rho, u);
            if (k <= cmax && (k - 2) % B == 0) ckpt[((k - 2) / B - 1) * stride] = real(u);
        }
    } else {
        zs.inv_z = inv(z);
        zs.inv_rho = 0.0;
        zs.start = start_index(n + 1, cabs(z));
        cx u = real(0.0);
        long long k = zs.start;
        // Down to the first checkpoint index (k = 2 mod B), then whole blocks k .. k - B + 1
        // whose first step gives a checkpoint, without branches, each redone on the checked
        // path in the rare lane whose arguments left the plain range.
        for (; k >= c0 && (k - 2) % B != 0; --k) u = step_complex((int)k, zs.inv_z, u);
        if constexpr (PLAIN) {
            for (; k > c0; k -= B) {
                cx first;
#pragma unroll
                for (int q = 0; q < B; ++q) {
                    u = step_complex_plain((int)(k - q), zs.inv_z, u, zs.pc);
                    if (q == 0) first = u;
                }
                if (k <= cmax) ckpt[((k - 2) / B - 1) * stride] = first;
            }
        }
        for (; k > c0; k -= B) {
            const cx u0 = u;
            PlainCheck pc;
            cx first;
#pragma unroll
            for (int q = 0; q < B; ++q) {
                u = step_complex_plain((int)(k - q), zs.inv_z, u, pc);
                if (q == 0) first = u;
            }
            if (!pc.ok()) {
                u = u0;
                for (int q = 0; q < B; ++q) {
                    u = step_complex((int)(k - q), zs.inv_z, u);
                    if (q == 0) first = u;
                }
            }
            if (k <= cmax) ckpt[((k - 2) / B - 1) * stride] = first;
        }
        if (k == c0) {
            u = step_complex((int)k, zs.inv_z, u);
            ckpt[0] = u;
        }
    }
    return zs;
}

// The real column starting `col` complex elements past a lane's complex slab pointer
// (slab_lane = slab + lane): a double column must be offset by the lane in doubles, not in
// complex elements, or lanes 32 apart would share slots.
__device__ __forceinline__ double *real_column(cx *slab_lane, int64_t col) {
    const int lane = threadIdx.x & 63;
    return reinterpret_cast<double *>(slab_lane - lane + col) + lane;
}

// One coefficient's effective log-derivative at a layered sphere's outer surface: as the
// correction delta to D_n(mx) or, near its poles, as H itself (scatrs/src/mie/kernel.rs:16-27).
struct OuterSide {
    cx va, vb;
    int direct_a, direct_b;
};

// The outer layer's quantities for a layered sphere's derivatives (mie/kernel.rs:29-38).
struct OuterDerivs {
    cx hp[2], hplus[2], extra_m[2];
};

// No outer sides: a homogeneous sphere.
struct NoOuter {
    static constexpr bool kLayered = false;
    __device__ OuterSide side(int) { return OuterSide{}; }
    __device__ OuterDerivs derivs(int) { return OuterDerivs{}; }
};

__device__ __forceinline__ double absorption_term(cx f, cx df, cx l, double n2) {
    return -((df.im - 2.0 * f.im * l.re) * n2);
}

// Whether a sink takes the branch-free blocks of orders (a sink with little state, which the
// block saves for its redo: kPlainBlocks = true).
template <class S, class = void>
struct plain_blocks : std::false_type {};
template <class S>
struct plain_blocks<S, std::void_t<decltype(S::kPlainBlocks)>>
    : std::integral_constant<bool, S::kPlainBlocks> {};

// Whether a sink wants a call at the end of each block of orders (block_end(k0, kend)).
template <class S, class = void>
struct has_block_end : std::false_type {};
template <class S>
struct has_block_end<S, std::void_t<decltype(&S::block_end)>> : std::true_type {};

// A range of orders k0..k1 of a sphere and the recurrence state on entry to order k0:
// psi = psi_{k0-1}(x), psi_prev = psi_{k0-2}(x), chi and chi_prev likewise (orders k0 <= n_s),
// and above n_s the ratio region's t, 1/xi and r of order k0 - 1.
struct OrderStart {
    int k0, k1;
    double psi, psi_prev, chi, chi_prev;
    cx t, inv_xi, r;
};

// The upward order loop (scatrs/src/mie/kernel.rs:325-559) after ratio_chain and x_chain.
// Calls sink.order(k, a, b) (or sink.order_d with the four derivatives) for k = 1..n and
// sink.done(abs_sum, jac). With a layered Outer source the per-order sides replace D_n(mx);
// with derivatives the sink then also receives, after order_d of the same order,
// sink.stored(k, inv_da, inv_db, g_a, g_b, d_m, md, outer), outer holding that order's
// inner-layer sensitivities. With `start` (an unblocked source only) the loop runs over
// start->k0..k1 from start's state, and abs_sum covers those orders.
template <bool DERIVS, class ZSrc, class Outer, class Sink>
__device__ __forceinline__ void mie_orders(double x, cx m, int n, const ZSrc &zs,
                                           const double *ux, int stride, Outer &outer,
                                           Sink &sink, const OrderStart *start = nullptr) {
    const cx z = m * x;
    const int n_s = upward_orders(x, n);
    const double inv_x = 1.0 / x;
    const cx inv_z = inv(z);
    const cx inv_m = inv(m);
    const cx m2 = m * m;
    const cx inv_m2 = inv_m * inv_m;
    const cx m2_minus_1 = (m - 1.0) * (m + 1.0);
    const cx c_a = -(m2_minus_1 * inv_m2);
    double sin_x, cos_x;
    sincos(x, &sin_x, &cos_x);
    double psi_prev = cos_x, psi = sin_x, chi_prev = -sin_x, chi = cos_x;
    cx t = real(0.0), inv_xi = real(0.0), r = real(0.0);
    if (n_s == 0) {
        t = real(sin_x) * inv(mk(sin_x, -cos_x));
        inv_xi = mk(sin_x, cos_x);
        r = mk(0.0, 1.0);
    }
    if (start) {
        psi = start->psi;
        psi_prev = start->psi_prev;
        chi = start->chi;
        chi_prev = start->chi_prev;
        t = start->t;
        inv_xi = start->inv_xi;
        r = start->r;
    }
    const bool m_is_one = (m.re == 1.0 && m.im == 0.0);
    double abs_sum = 0.0;
    double jac0 = 0.0, jac1 = 0.0, jac2 = 0.0;
    // One order of the loop, with u_{k+1}(mx) in registers. The plain forms (for spheres
    // without derivatives) take their branch without testing for it and invert on the plain
    // path of inv, noting the paired inverses' arguments in pc and the single ones' in pcr:
    // mode 1 for orders up to n_s (the upward branch), mode 2 for orders above it (the ratio
    // region); mode 0 is the checked form of either.
    auto order_step_t = [&](int k, const cx u_z1, auto mode, PlainCheck &pc, PlainCheck &pcr) {
        constexpr int MODE = decltype(mode)::value;
        constexpr bool FAST = MODE == 1;
        const bool upward = MODE == 1 ? true : (MODE == 2 ? false : k <= n_s);
        const double kf = (double)k;
        const double k_x = kf * inv_x;

        const cx d = inv_z * (double)(k + 1) - u_z1;
        OuterSide side{};
        cx hh_a = d, hh_b = d;
        if constexpr (Outer::kLayered) {
            side = outer.side(k);
            hh_a = side.direct_a ? side.va : d + side.va;
            hh_b = side.direct_b ? side.vb : d + side.vb;
        }
        const cx d_m = hh_a * inv_m;
        const cx md = hh_b * m;
        const cx big_a = d_m + k_x;
        const cx big_b = md + k_x;
        cx an, bn, inv_da, inv_db, g_a, g_b, h_a, h_b;
        double na = 0.0, nb = 0.0;
        if (upward) {
            const double c = (double)(2 * k - 1) * inv_x;
            const double psi_next = c * psi - psi_prev;
            psi_prev = psi;
            psi = psi_next;
            const double chi_next = c * chi - chi_prev;
            chi_prev = chi;
            chi = chi_next;
            const cx xi = mk(psi, -chi);
            const cx xi_prev = mk(psi_prev, -chi_prev);
            if constexpr (FAST) {
                inv_plain_pair(big_a * xi - xi_prev, big_b * xi - xi_prev, pc, inv_da, inv_db,
                               na, nb);
            } else {
                (void)pc;
                (void)pcr;
                inv_da = inv(big_a * xi - xi_prev);
                inv_db = inv(big_b * xi - xi_prev);
            }
            if constexpr (Outer::kLayered) {
                if (!side.direct_a) {
                    const cx own = m_is_one ? real(0.0) : ((d * inv_m + k_x) * psi - psi_prev);
                    an = (own + (side.va * inv_m) * psi) * inv_da;
                } else {
                    an = (big_a * psi - psi_prev) * inv_da;
                }
                if (!side.direct_b) {
                    const cx own = m_is_one ? real(0.0) : ((d * m + k_x) * psi - psi_prev);
                    bn = (own + (m * side.vb) * psi) * inv_db;
                } else {
                    bn = (big_b * psi - psi_prev) * inv_db;
                }
            } else {
                (void)m_is_one;
                an = (big_a * psi - psi_prev) * inv_da;
                bn = (big_b * psi - psi_prev) * inv_db;
            }
            g_a = xi * inv_da;
            g_b = xi * inv_db;
            h_a = xi_prev * inv_da;
            h_b = xi_prev * inv_db;
            if (MODE == 0 && k == n_s) {
                inv_xi = inv(xi);
                t = real(psi) * inv_xi;
                r = xi_prev * inv_xi;
            }
        } else {
            if constexpr (MODE == 2)
                r = inv_plain((double)(2 * k - 1) * inv_x - r, pcr);
            else
                r = inv((double)(2 * k - 1) * inv_x - r);
            t = (t * r) * ux[k * stride];
            inv_xi = inv_xi * r;
            const double ux1 = ux[(k + 1) * stride];
            cx num_a = c_a * ((double)(k + 1) * inv_x) + ux1 - u_z1 * inv_m;
            cx num_b = ux1 - m * u_z1;
            if constexpr (Outer::kLayered) {
                const cx d_x = real((double)(k + 1) * inv_x - ux1);
                num_a = side.direct_a ? d_m - d_x : num_a + side.va * inv_m;
                num_b = side.direct_b ? md - d_x : num_b + m * side.vb;
            }
            cx ia, ib;
            if constexpr (MODE == 2) {
                double ta, tb;
                inv_plain_pair(big_a - r, big_b - r, pc, ia, ib, ta, tb);
            } else {
                ia = inv(big_a - r);
                ib = inv(big_b - r);
            }
            an = (t * num_a) * ia;
            bn = (t * num_b) * ib;
            inv_da = inv_xi * ia;
            inv_db = inv_xi * ib;
            g_a = ia;
            g_b = ib;
            h_a = r * ia;
            h_b = r * ib;
        }
        // |1/Delta|^2: from the paired reciprocal on the plain upward path
        if constexpr (!FAST) {
            na = norm2(inv_da);
            nb = norm2(inv_db);
        }
        const double weight = (double)(2 * k + 1);
        abs_sum += weight * (-(d_m.im * na) - md.im * nb);
        if constexpr (!DERIVS) {
            sink.order(k, an, bn);
        } else {
            const double kk1 = kf * (double)(k + 1);
            const cx ida2 = inv_da * inv_da, idb2 = inv_db * inv_db;
            const double k_x2 = k_x * inv_x;
            cx da_dm, db_dm, da_dx, db_dx, fa_m, fb_m, dpa, dpb;
            if constexpr (!Outer::kLayered) {
                const cx one_u2 = 1.0 + u_z1 * u_z1;
                const cx dp = ((u_z1 * 2.0) * inv_z - inv_z * inv_z) * (double)(k + 1) - one_u2;
                const cx d_plus = u_z1 * weight - z * one_u2;
                const cx d_minus = d * 2.0 - d_plus;
                da_dm = (times_i(d_minus) * inv_m2) * ida2;
                db_dm = -(times_i(d_plus) * idb2);
                da_dx = ((times_i(m2_minus_1) * (d * d + kk1 * inv_x * inv_x)) * inv_m2) * ida2;
                db_dx = times_i(m2_minus_1) * idb2;
                fa_m = -(d_minus * inv_m2);
                fb_m = d_plus;
                dpa = dp;
                dpb = dp;
            } else {
                const OuterDerivs od = outer.derivs(k);
                dpa = od.hp[0];
                dpb = od.hp[1];
                const cx minus_a = hh_a * 2.0 - od.hplus[0];
                fa_m = -(minus_a * inv_m2) + od.extra_m[0] * inv_m;
                fb_m = od.hplus[1] + m * od.extra_m[1];
                da_dm = -(times_i(fa_m) * ida2);
                db_dm = -(times_i(fb_m) * idb2);
                da_dx = ((times_i(m2_minus_1) * (hh_a * hh_a + kk1 * inv_x * inv_x)) * inv_m2) * ida2;
                db_dx = times_i(m2_minus_1) * idb2;
            }
            const cx la_x = g_a * (dpa - k_x2 - big_a * k_x + 1.0) + h_a * d_m;
            const cx lb_x = g_b * (m2 * dpb - k_x2 - big_b * k_x + 1.0) + h_b * md;
            const cx la_m = fa_m * g_a, lb_m = fb_m * g_b;
            jac0 += weight * (absorption_term(d_m, dpa, la_x, na) +
                              absorption_term(md, m2 * dpb, lb_x, nb));
            jac1 += weight * (absorption_term(d_m, fa_m, la_m, na) +
                              absorption_term(md, fb_m, lb_m, nb));
            jac2 += weight * (absorption_term(d_m, times_i(fa_m), times_i(la_m), na) +
                              absorption_term(md, times_i(fb_m), times_i(lb_m), nb));
            sink.order_d(k, an, bn, da_dx, da_dm, db_dx, db_dm);
            if constexpr (Outer::kLayered) sink.stored(k, inv_da, inv_db, g_a, g_b, d_m, md, outer);
        }
    };
    auto order_step = [&](int k, const cx u_z1) {
        PlainCheck pc;
        order_step_t(k, u_z1, std::integral_constant<int, 0>{}, pc, pc);
    };
    if constexpr (ZSrc::kBlocked) {
        constexpr int B = ZSrc::kBlock;
        for (int j = 0; 1 + j * B <= n; ++j) {
            zs.regen(j);
            const int k0 = 1 + j * B;
            const int kend = n < k0 + B - 1 ? n : k0 + B - 1;
            if constexpr (!DERIVS && !Outer::kLayered &&
                          (ZSrc::kPlain || plain_blocks<Sink>::value)) {
                // A block wholly below n_s: B orders without branches, redone on the checked
                // path (from the saved state) in the rare lane whose inverses left the plain
                // range.
                if constexpr (ZSrc::kPlain) {
                    // Unrolled 8 orders at a time (blocks of 16 in two halves).
                    constexpr int U = B < 8 ? B : 8;
                    if (k0 + B - 1 < n_s) {
#pragma unroll 1
                        for (int h = 0; h < B; h += U) {
#pragma unroll
                            for (int q = h; q < h + U; ++q)
                                order_step_t(k0 + q, zs.lds[q * 64],
                                             std::integral_constant<int, 1>{}, zs.pc_pair, zs.pc);
                        }
                        if constexpr (has_block_end<Sink>::value) sink.block_end(k0, k0 + B - 1);
                        continue;
                    }
                    // A full block above n_s (the ratio region).
                    if (k0 > n_s && kend == k0 + B - 1) {
#pragma unroll 1
                        for (int h = 0; h < B; h += U) {
#pragma unroll
                            for (int q = h; q < h + U; ++q)
                                order_step_t(k0 + q, zs.lds[q * 64],
                                             std::integral_constant<int, 2>{}, zs.pc_pair, zs.pc);
                        }
                        if constexpr (has_block_end<Sink>::value) sink.block_end(k0, kend);
                        continue;
                    }
                } else if (k0 + B - 1 < n_s) {
                    const double s_psi = psi, s_psi_prev = psi_prev, s_chi = chi,
                                 s_chi_prev = chi_prev, s_abs = abs_sum;
                    const Sink s_sink = sink;
                    PlainCheck pc;
#pragma unroll
                    for (int q = 0; q < B; ++q)
                        order_step_t(k0 + q, zs.lds[q * 64], std::integral_constant<int, 1>{}, pc,
                                     pc);
                    if (!pc.ok_pair()) {
                        psi = s_psi;
                        psi_prev = s_psi_prev;
                        chi = s_chi;
                        chi_prev = s_chi_prev;
                        abs_sum = s_abs;
                        sink = s_sink;
                        for (int q = 0; q < B; ++q) order_step(k0 + q, zs.lds[q * 64]);
                    }
                    if constexpr (has_block_end<Sink>::value) sink.block_end(k0, k0 + B - 1);
                    continue;
                }
            }
#pragma unroll 1
            for (int k = k0; k <= kend; ++k) order_step(k, zs.lds[(k - k0) * 64]);
            // A sink may use the block's tile slots, free once their order is done.
            if constexpr (has_block_end<Sink>::value) sink.block_end(k0, kend);
        }
    } else {
        // Read two orders ahead with one prefetch buffer per order of a pair, so every load
        // has a whole order of arithmetic to arrive and stays in its registers.
        const cx *uz = zs.uz;
        const int zst = zs.stride;
        const int ka = start ? start->k0 : 1, kb = start ? start->k1 : n;
        cx pa = ka <= kb ? uz[(ka + 1) * zst] : real(0.0);
        cx pb = ka + 1 <= kb ? uz[(ka + 2) * zst] : real(0.0);
        int k = ka;
        for (; k + 1 <= kb; k += 2) {
            order_step(k, pa);
            if (k + 3 <= kb + 1) pa = uz[(k + 3) * zst];
            order_step(k + 1, pb);
            if (k + 4 <= kb + 1) pb = uz[(k + 4) * zst];
        }
        if (k == kb) order_step(k, pa);
    }
    sink.done(abs_sum, jac0, jac1, jac2);
}

// The derivative sums of the efficiency series along one particle parameter group (x, Re m,
// Im m of one layer; efficiencies.rs:89-146): order k adds its coefficients' derivatives
// dak[p], dbk[p] with w = 2k + 1, own = (2k + 1)/(k (k + 1)), pair = (k - 1)(k + 1)/k and the
// previous order's coefficients a_prev, b_prev.
struct DerivSums {
    double d_sca[3] = {0.0, 0.0, 0.0}, d_g[3] = {0.0, 0.0, 0.0};
    double jac_abs[3] = {0.0, 0.0, 0.0};
    cx d_back[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    cx da_prev[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};
    cx db_prev[3] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}};

    __device__ void add(int k, double w, double own, double pair, cx a, cx b, cx a_prev,
                        cx b_prev, const cx dak[3], const cx dbk[3]) {
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            d_sca[p] += w * 2.0 * (re_mul_conj(dak[p], a) + re_mul_conj(dbk[p], b));
            const cx t = (dak[p] - dbk[p]) * w;
            d_back[p] = (k % 2 == 0) ? d_back[p] + t : d_back[p] - t;
            if (k > 1) {
                d_g[p] += pair * (re_mul_conj(da_prev[p], a) + re_mul_conj(a_prev, dak[p]) +
                                  re_mul_conj(db_prev[p], b) + re_mul_conj(b_prev, dbk[p]));
            }
            d_g[p] += own * (re_mul_conj(dak[p], b) + re_mul_conj(a, dbk[p]));
            da_prev[p] = dak[p];
            db_prev[p] = dbk[p];
        }
    }

    // The Jacobian [d/dx, d/dRe m, d/dIm m][Q_ext..g] (efficiencies.rs:192-230) from the values
    // q and the value sums sca, back; `outer` adds the 1/x^2 prefactor's term to d/dx (the
    // outer layer's size parameter sets the efficiencies' normalization).
    __device__ void jacobian(double x, const double q[5], double sca, cx back, bool outer,
                             double out[15]) const {
#pragma clang fp contract(off)
        const double inv_x = 1.0 / x, inv_x2 = inv_x * inv_x, two_x2 = 2.0 * inv_x2;
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            const double xt = (p == 0 && outer) ? 2.0 * inv_x : 0.0;
            const double q_sca = two_x2 * d_sca[p] - xt * q[1];
            const double q_abs = two_x2 * jac_abs[p] - xt * q[2];
            const double q_back = 2.0 * re_mul_conj(d_back[p], back) * inv_x2 - xt * q[3];
            const double gg = sca == 0.0 ? 0.0 : (2.0 * d_g[p] - q[4] * d_sca[p]) / sca;
            out[p * 5 + 0] = q_sca + q_abs;
            out[p * 5 + 1] = q_sca;
            out[p * 5 + 2] = q_abs;
            out[p * 5 + 3] = q_back;
            out[p * 5 + 4] = gg;
        }
    }
};

// The efficiency series (scatrs/src/mie/efficiencies.rs:64-187) accumulated order by order:
// Q_sca = (2/x^2) sca, Q_back = |back|^2 / x^2, g = 2 G / sca, and with derivatives the sums
// of efficiencies.rs:89-146 for d/dx, d/dRe m, d/dIm m.
template <bool DERIVS>
struct EfficiencySink {
    static constexpr bool kPlainBlocks = !DERIVS;
    double sca = 0.0, g = 0.0, abs_sum = 0.0;
    cx back = {0.0, 0.0};
    cx a_prev = {0.0, 0.0}, b_prev = {0.0, 0.0};
    DerivSums ds;

    // 1/k of the coming order: each order takes one reciprocal, 1/(k + 1), and keeps it for the
    // next, for its pair and own weights pair(k - 1) = (k - 1)(k + 1)/k and
    // own(k) = (2k + 1)/(k (k + 1)).
    double rk = 1.0;

    // The weights of order k: own(k) and pair(k - 1), and 1/(k + 1).
    __device__ void weights(int k, double *own, double *pair, double *rk1) const {
        const double w = (double)(2 * k + 1), kf = (double)k;
        *rk1 = recip(kf + 1.0);
        *own = w * (rk * *rk1);
        const double km = kf - 1.0;
        *pair = (km * (km + 2.0)) * rk;
    }

    __device__ void order(int k, cx a, cx b) {
        const double w = (double)(2 * k + 1);
        double own, pair, rk1;
        weights(k, &own, &pair, &rk1);
        sca += w * (norm2(a) + norm2(b));
        const cx t = (a - b) * w;
        back = (k % 2 == 0) ? back + t : back - t;
        if (k > 1) g += pair * (re_mul_conj(a_prev, a) + re_mul_conj(b_prev, b));
        g += own * re_mul_conj(a, b);
        a_prev = a;
        b_prev = b;
        rk = rk1;
    }

    __device__ void order_d(int k, cx a, cx b, cx da_dx, cx da_dm, cx db_dx, cx db_dm) {
        const cx dak[3] = {da_dx, da_dm, times_i(da_dm)};
        const cx dbk[3] = {db_dx, db_dm, times_i(db_dm)};
        double own, pair, rk1;
        weights(k, &own, &pair, &rk1);
        ds.add(k, (double)(2 * k + 1), own, pair, a, b, a_prev, b_prev, dak, dbk);
        order(k, a, b);
    }

    template <class O>
    __device__ void stored(int, cx, cx, cx, cx, cx, cx, const O &) {}

    // A lossless layered sphere absorbs nothing, and stays lossless as x or Re m changes
    // (layered/kernel.rs:453-464).
    __device__ void zero_absorption() {
        abs_sum = 0.0;
        ds.jac_abs[0] = 0.0;
        ds.jac_abs[1] = 0.0;
    }

    __device__ void done(double abs_s, double j0, double j1, double j2) {
        abs_sum = abs_s;
        ds.jac_abs[0] = j0;
        ds.jac_abs[1] = j1;
        ds.jac_abs[2] = j2;
    }

    // [Q_ext, Q_sca, Q_abs, Q_back, g] (efficiencies.rs:169-187).
    __device__ void values(double x, double q[5]) const {
        // Q_ext is Q_sca + Q_abs exactly, as the reference forms it: no contraction here.
#pragma clang fp contract(off)
        const double inv_x = 1.0 / x, inv_x2 = inv_x * inv_x, two_x2 = 2.0 * inv_x2;
        const double q_sca = two_x2 * sca, q_abs = two_x2 * abs_sum;
        q[0] = q_sca + q_abs;
        q[1] = q_sca;
        q[2] = q_abs;
        q[3] = norm2(back) * inv_x2;
        q[4] = sca == 0.0 ? 0.0 : 2.0 * g / sca;
    }

    // The Jacobian [d/dx, d/dRe m, d/dIm m][Q_ext..g] (efficiencies.rs:192-230).
    __device__ void jacobian(double x, const double q[5], bool outer, double out[15]) const {
        ds.jacobian(x, q, sca, back, outer, out);
    }
};

// The sink of a sphere's checked redo: its orders below floor(x) take the checked path one by
// one, without the branch-free blocks.
template <bool DERIVS>
struct CheckedEfficiencySink : EfficiencySink<DERIVS> {
    static constexpr bool kPlainBlocks = false;
};

// One sphere's efficiencies (and Jacobian): the downward u_k(mx) pass keeps one checkpoint per
// kSphereBlock indices in ck (lane-strided by 64), the order loop regenerates each block into
// the lane's LDS tile, u_k(x) above floor(x) lives in ux. Without derivatives (and with
// `plain`) the whole sphere first runs on the plain path (no per-block state kept for a redo)
// and, in the rare lane whose inverses left the plain range, again on the checked path.
template <int B, bool DERIVS>
__device__ __forceinline__ void sphere_efficiencies(double x, cx m, int n, cx *ck, double *ux,
                                                    cx *tile, bool plain, double q[5],
                                                    double *jac15) {
    x_chain(x, n, upward_orders(x, n), ux, 64);
    NoOuter none;
    if constexpr (!DERIVS) {
        (void)jac15;
        if (plain) {
            EfficiencySink<false> s;
            const auto zs = ratio_chain_ckpt<B, true>(m, x, n, ck, 64, tile);
            mie_orders<false>(x, m, n, zs, ux, 64, none, s);
            if (zs.plain_ok()) {
                s.values(x, q);
                return;
            }
        }
        CheckedEfficiencySink<false> c;
        const auto zc = ratio_chain_ckpt<B>(m, x, n, ck, 64, tile);
        (void)plain;
        mie_orders<false>(x, m, n, zc, ux, 64, none, c);
        c.values(x, q);
    } else {
        EfficiencySink<true> s;
        const auto zs = ratio_chain_ckpt<B>(m, x, n, ck, 64, tile);
        mie_orders<true>(x, m, n, zs, ux, 64, none, s);
        s.values(x, q);
        s.jacobian(x, q, true, jac15);
    }
}

}  // namespace scatrs
