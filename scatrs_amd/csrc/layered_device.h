// Per-lane layered-sphere solver for gfx950 (scatrs/src/layered/kernel.rs:28-587): the
// effective log-derivatives H_a, H_b from the core outward in the pole-free form, fed order by
// order into the homogeneous outer step of mie_orders. Unlike the reference, which runs the
// layer recursion over every order first and stores the per-order sides, the sweep here runs
// lazily inside the outer step's order loop (one order of recursion per outer order), so
// nothing per order is stored beyond the ratio columns.
//
// Ratio columns in the lane's scratch slab (column c at slab + c * cs, element k at k * 64):
//   0: u_k(m_{L-1} x_{L-1}) (the outer layer's, as a sphere's), 1: u_k(m_0 x_0) (the core),
//   2l: u_k(m_l x_{l-1}) and 2l + 1: u_k(m_l x_l) (l < L - 1), 2L - 1: u_k(x) (real).
#pragma once

#include "mie_device.h"

namespace scatrs {

constexpr int kMaxLayersDevice = 8;

template <int L>
struct LayeredColumns {
    static constexpr int kComplex = L == 1 ? 1 : 2 * L - 1;
};

// Pi_n(z) = psi_n zeta_n and w_{n+1} = zeta_{n+1}/zeta_n of one argument (layered/kernel.rs:28-97).
struct ArgState {
    cx z, inv_z, pi, w, w_n;
};

__device__ inline cx two_i(cx z) { return mk(-(z.im * 2.0), z.re * 2.0); }

// Both forms are evaluated and one is selected per lane, so the lanes of a wavefront that
// take different forms run one instruction stream (the unused form may overflow harmlessly).
__device__ inline ArgState arg_new(cx z, bool chi) {
    ArgState a;
    a.z = z;
    a.inv_z = inv(z);
    const cx s = csin(z), c = ccos(z);
    const cx pi_chi = s * c, w_chi = a.inv_z + s * inv(c);
    const cx pi_xi = -(exp_m1(two_i(z)) * 0.5), w_xi = a.inv_z - mk(0.0, 1.0);
    a.pi = chi ? pi_chi : pi_xi;
    a.w = chi ? w_chi : w_xi;
    a.w_n = real(0.0);
    return a;
}

__device__ __forceinline__ void arg_step(ArgState &a, int n, cx u_n) {
    a.w_n = a.w;
    a.pi = (a.pi * u_n) * a.w;
    a.w = a.inv_z * (double)(2 * n + 1) - inv(a.w);
}

__device__ __forceinline__ cx arg_d1_prime(const ArgState &a, int n, cx u) {
    const cx iz = a.inv_z;
    return ((u * 2.0) * iz - iz * iz) * (double)(n + 1) - (1.0 + u * u);
}

__device__ __forceinline__ cx arg_d1_plus(const ArgState &a, int n, cx u) {
    return u * (double)(2 * n + 1) - a.z * (1.0 + u * u);
}

// Sensitivities of one layer's H at one order (layered/kernel.rs:127-146).
struct Sens {
    cx hv, r2, em, hp, hplus;
};

struct LStep {
    cx u_prev, u1, d1i, d3i, d1o, d3o, pio, p, w, u2, s;
    int n;
    const ArgState *a1, *a2;
};

// H at a layer's outer surface from H at its inner surface (layered/kernel.rs:148-232).
template <bool DERIVS>
__device__ inline cx side_step(cx mi, cx mo, bool prev_direct, cx prev_val, cx h, cx lead,
                               const LStep &s, bool lead_is_a, bool *direct, cx *val,
                               Sens *sens) {
    const cx mh = mi * h;
    const cx g1 = prev_direct ? mh - mo * s.d1i : ((lead - mi * s.u_prev) + mo * s.u1) + mi * prev_val;
    const cx g2 = mh - mo * s.d3i;
    const cx inv_den = inv(s.pio * g2 - s.p * g1);
    const cx delta = -((((s.w * s.p) * g1) * inv_den) * inv(s.pio));
    const cx hd = s.d1o + delta;
    cx h_out;
    const double lim = fmax(norm1(s.d1o), norm1(delta));
    if (norm1(hd) * 4.0 >= lim) {
        *direct = false;
        *val = delta;
        h_out = hd;
    } else {
        const cx hp = s.d3o - (s.w * g2) * inv_den;
        *direct = true;
        *val = hp;
        h_out = hp;
    }
    if constexpr (DERIVS) {
        const cx inv_mo = inv(mo);
        const cx t = mh * inv_mo;
        const cx r = -(((mo * s.w) * s.s) * inv_den);
        const cx e = g1 * inv_mo;
        const double sigma = lead_is_a ? 1.0 : -1.0;
        const cx d1 = s.d1i;
        const cx z1 = s.a1->z;
        const cx own = lead_is_a ? d1 * 2.0 - arg_d1_plus(*s.a1, s.n, s.u1)
                                 : -arg_d1_plus(*s.a1, s.n, s.u1);
        cx em;
        if (norm1(t) * 4.0 >= fmax(norm1(d1), norm1(e))) {
            em = own + e * (sigma + z1 * (d1 * 2.0 + e));
        } else {
            const cx iz = s.a1->inv_z;
            const cx v = ((real((double)s.n * (double)(s.n + 1)) * iz) * iz) - 1.0;
            em = t * sigma - z1 * (v - t * t);
        }
        cx hp, hplus;
        if (!*direct) {
            const cx d2 = s.d1o * 2.0 + delta;
            hp = arg_d1_prime(*s.a2, s.n, s.u2) - delta * d2;
            hplus = arg_d1_plus(*s.a2, s.n, s.u2) + delta * (1.0 - s.a2->z * d2);
        } else {
            const cx hv = *val;
            const cx iz = s.a2->inv_z;
            const cx v = ((real((double)s.n * (double)(s.n + 1)) * iz) * iz) - 1.0;
            hp = v - hv * hv;
            hplus = hv + s.a2->z * hp;
        }
        sens->hv = h_out;
        sens->r2 = r * r;
        sens->em = em;
        sens->hp = hp;
        sens->hplus = hplus;
    }
    return h_out;
}

// The outer-side source of a layered sphere for mie_orders: side(k) advances the layer
// recursion by one order; derivs(k) returns the outer layer's derivative quantities of the
// same order, and dh holds dH_a/dx_l, dH_a/dm_l, dH_b/dx_l, dH_b/dm_l of the inner layers.
template <int L, bool DERIVS>
struct LayeredOuter {
    static constexpr bool kLayered = true;
    static constexpr int NL = L - 1;
    cx m[L];
    double xl[L];
    const cx *slab;
    int64_t cs;
    ArgState in[NL], out[NL];
    cx p[NL], lead[NL], sv[NL];
    bool chi[NL];
    cx core_inv_z, core_z;
    OuterDerivs od;
    cx dh[NL][4];

    __device__ cx col(int c, int k) const { return slab[c * cs + (int64_t)k * 64]; }
    // u index j of the reference's layout (0: core, 2l-1: m_l x_{l-1}, 2l: m_l x_l) -> column
    __device__ cx U(int j, int k) const { return col(j + 1, k); }
    __device__ cx U2(int l, int k) const { return l == L - 1 ? col(0, k) : U(2 * l, k); }

    __device__ void init() {
        for (int l = 1; l < L; ++l) {
            const bool c = m[l].im * xl[l] < 0.5;
            chi[l - 1] = c;
            const cx aa = m[l], bb = m[l - 1];
            lead[l - 1] = ((aa - bb) * (aa + bb)) * inv(aa * bb);
            const cx z1 = m[l] * xl[l - 1], z2 = m[l] * xl[l];
            in[l - 1] = arg_new(z1, c);
            out[l - 1] = arg_new(z2, c);
            // chi and xi forms selected per lane, as in arg_new
            const cx c1 = ccos(z1), c2 = ccos(z2);
            const double dx = xl[l] - xl[l - 1];
            const double ea = -(m[l].im * dx * 2.0), eb = m[l].re * dx * 2.0;
            const cx e = mk(cos(eb), sin(eb)) * exp(ea);
            const cx p_chi = ((csin(z1) * inv(c1)) * c2) * c2;
            const cx p_xi = -((e * exp_m1(two_i(z1))) * 0.5);
            p[l - 1] = c ? p_chi : p_xi;
            if constexpr (DERIVS) {
                const double b2 = m[l].re * dx;
                const cx s_chi = c2 * inv(c1);
                const cx s_xi = mk(cos(b2), sin(b2)) * exp(-(m[l].im * dx));
                sv[l - 1] = c ? s_chi : s_xi;
            }
        }
        core_z = m[0] * xl[0];
        core_inv_z = inv(core_z);
    }

    __device__ OuterSide side(int k) {
#pragma unroll
        for (int l = 1; l < L; ++l) {
            const cx u1k = U(2 * l - 1, k), u2k = U2(l, k);
            arg_step(in[l - 1], k, u1k);
            arg_step(out[l - 1], k, u2k);
            p[l - 1] = (((p[l - 1] * u1k) * inv(in[l - 1].w_n)) * out[l - 1].w_n) * out[l - 1].w_n;
            if constexpr (DERIVS) sv[l - 1] = (sv[l - 1] * out[l - 1].w_n) * inv(in[l - 1].w_n);
        }
        const cx u0 = U(0, k + 1);
        cx ha = core_inv_z * (double)(k + 1) - u0;
        cx hb = ha;
        bool da = false, db = false;
        cx va = real(0.0), vb = real(0.0);
        Sens sa[NL], sb[NL];
#pragma unroll
        for (int l = 1; l < L; ++l) {
            const ArgState &ai = in[l - 1], &ao = out[l - 1];
            LStep s;
            s.u_prev = (l == 1) ? u0 : U(2 * (l - 1), k + 1);
            s.u1 = U(2 * l - 1, k + 1);
            s.u2 = U2(l, k + 1);
            s.d1i = ai.inv_z * (double)(k + 1) - s.u1;
            s.d3i = ai.inv_z * (double)(k + 1) - ai.w;
            s.d1o = ao.inv_z * (double)(k + 1) - s.u2;
            s.d3o = ao.inv_z * (double)(k + 1) - ao.w;
            s.pio = ao.pi;
            s.p = p[l - 1];
            s.w = chi[l - 1] ? real(-1.0) : mk(0.0, 1.0);
            s.s = DERIVS ? sv[l - 1] : real(0.0);
            s.n = k;
            s.a1 = &ai;
            s.a2 = &ao;
            const cx ld = lead[l - 1] * ((double)(k + 1) / xl[l - 1]);
            bool nda, ndb;
            cx nva, nvb;
            const cx ah = side_step<DERIVS>(m[l], m[l - 1], da, va, ha, ld, s, true, &nda, &nva,
                                            &sa[l - 1]);
            const cx bh = side_step<DERIVS>(m[l - 1], m[l], db, vb, hb, real(0.0), s, false,
                                            &ndb, &nvb, &sb[l - 1]);
            ha = ah;
            hb = bh;
            da = nda;
            va = nva;
            db = ndb;
            vb = nvb;
        }
        if constexpr (DERIVS) {
            // The adjoint sweep over the layers (layered/kernel.rs:479-543).
            const double nn1 = (double)k * (double)(k + 1);
            const cx c_h = core_inv_z * (double)(k + 1) - u0;
            const cx c_hplus = u0 * (double)(2 * k + 1) - core_z * (1.0 + u0 * u0);
            cx extra[2];
#pragma unroll
            for (int sd = 0; sd < 2; ++sd) {
                cx dxv[L], dmv[L];
#pragma unroll
                for (int l = 0; l < L; ++l) dxv[l] = dmv[l] = real(0.0);
                cx lambda = real(1.0);
#pragma unroll
                for (int l = L - 1; l >= 1; --l) {
                    const Sens e = sd == 0 ? sa[l - 1] : sb[l - 1];
                    const cx ml = m[l], mp = m[l - 1];
                    dmv[l] = dmv[l] + ((lambda * e.r2) * e.em) * inv(ml);
                    cx h, hplus;
                    if (l == 1) {
                        h = c_h;
                        hplus = c_hplus;
                    } else {
                        const Sens pp = sd == 0 ? sa[l - 2] : sb[l - 2];
                        h = pp.hv;
                        hplus = pp.hplus;
                    }
                    const cx c = (ml - mp) * (ml + mp);
                    const cx w = lambda * e.r2;
                    const cx inv_mp = inv(mp);
                    cx jump, kk;
                    if (sd == 0) {
                        const double inv_xl = 1.0 / xl[l - 1];
                        jump = (ml * c) * (real(nn1 * inv_xl * inv_xl) * inv(((mp * mp) * ml) * ml) +
                                           ((h * h) * inv_mp) * inv_mp);
                        kk = -(((ml * inv_mp) * inv_mp) * (h * 2.0 - hplus));
                    } else {
                        const cx inv_ml = inv(ml);
                        jump = c * inv_ml;
                        kk = hplus * inv_ml;
                    }
                    dxv[l - 1] = dxv[l - 1] + w * jump;
                    dmv[l - 1] = dmv[l - 1] + w * kk;
                    const cx mi = sd == 0 ? ml : mp, mo = sd == 0 ? mp : ml;
                    lambda = (w * mi) * inv(mo);
                }
                extra[sd] = dmv[L - 1];
#pragma unroll
                for (int l = 0; l < L - 1; ++l) {
                    dh[l][2 * sd] = dxv[l];
                    dh[l][2 * sd + 1] = dmv[l];
                }
            }
            od.hp[0] = sa[L - 2].hp;
            od.hp[1] = sb[L - 2].hp;
            od.hplus[0] = sa[L - 2].hplus;
            od.hplus[1] = sb[L - 2].hplus;
            od.extra_m[0] = extra[0];
            od.extra_m[1] = extra[1];
        }
        OuterSide o;
        o.direct_a = da;
        o.direct_b = db;
        o.va = va;
        o.vb = vb;
        return o;
    }

    __device__ OuterDerivs derivs(int) { return od; }
};

// Solves one layered sphere of interface size parameters xl[0..L-1] (the outer one last;
// indices mm interleaved re, im per layer) with n orders, feeding sink as mie_orders does. The
// columns of the lane's slab are filled first.
template <int L, bool DERIVS, class Sink>
__device__ inline void layered_solve_xl(const double *xl, const double *mm, int n, cx *slab,
                                        int64_t rows, Sink &sink,
                                        LayeredOuter<L, DERIVS> *keep = nullptr) {
    const int64_t cs = rows * 64;
    LayeredOuter<L, DERIVS> o;
#pragma unroll
    for (int l = 0; l < L; ++l) {
        o.xl[l] = xl[l];
        o.m[l] = mk(mm[2 * l], mm[2 * l + 1]);
    }
    const double x = o.xl[L - 1];
    o.slab = slab;
    o.cs = cs;
    double *ux = real_column(slab, (2 * L - 1) * cs);
    ratio_chain(o.m[L - 1], x, n, slab, 64);
    x_chain(x, n, upward_orders(x, n), ux, 64);
    ratio_chain(o.m[0], o.xl[0], n, slab + cs, 64);
#pragma unroll
    for (int l = 1; l < L; ++l) {
        ratio_chain(o.m[l], o.xl[l - 1], n, slab + 2 * l * cs, 64);
        if (l < L - 1) ratio_chain(o.m[l], o.xl[l], n, slab + (2 * l + 1) * cs, 64);
    }
    o.init();
    const ZColumn zs{slab, 64};
    mie_orders<DERIVS>(x, o.m[L - 1], n, zs, ux, 64, o, sink);
    bool lossless = true;
#pragma unroll
    for (int l = 0; l < L; ++l) lossless = lossless && o.m[l].im == 0.0;
    if (lossless) sink.zero_absorption();
    if (keep) *keep = o;
}

// Solves one layered sphere of outer size parameter x (x_l = fr[l] x, indices mm interleaved
// re, im per layer) with n orders, as layered_solve_xl.
template <int L, bool DERIVS, class Sink>
__device__ inline void layered_solve(double x, const double *fr, const double *mm, int n,
                                     cx *slab, int64_t rows, Sink &sink,
                                     LayeredOuter<L, DERIVS> *keep = nullptr) {
    double xl[L];
#pragma unroll
    for (int l = 0; l < L; ++l) xl[l] = l < L - 1 ? fr[l] * x : x;
    layered_solve_xl<L, DERIVS>(xl, mm, n, slab, rows, sink, keep);
}

}  // namespace scatrs
