/* CPU FP64 oracle: a plain-C restatement of the reference's algorithms for the
 * size-distribution ensemble path. TEST INFRASTRUCTURE ONLY (see scatrs_oracle.h): it checks
 * and times the GPU library and is never part of it.
 *
 * Compiled with -ffp-contract=off so that every operation rounds once, as the reference's Rust
 * does (no fused multiply-adds), and complex products as num-complex 0.4.6 forms them:
 * (a + ib)(c + id) = (ac - bd) + i(ad + bc). */
#include "scatrs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef orc_cx cx;

static const double EPS = 2.220446049250313080847e-16; /* f64::EPSILON = 2^-52 */
static const double PI = 3.14159265358979323846264338327950288;

/* ---- complex arithmetic (complex.rs:21-56 and num-complex's operators) ---- */
static inline cx C(double re, double im) { cx z = {re, im}; return z; }
static inline cx R(double re) { return C(re, 0.0); }
static inline cx add(cx a, cx b) { return C(a.re + b.re, a.im + b.im); }
static inline cx sub(cx a, cx b) { return C(a.re - b.re, a.im - b.im); }
static inline cx mul(cx a, cx b) { return C(a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re); }
static inline cx scl(cx a, double t) { return C(a.re * t, a.im * t); }
static inline cx neg(cx a) { return C(-a.re, -a.im); }
static inline cx times_i(cx u) { return C(-u.im, u.re); }
static inline double norm2(cx a) { return a.re * a.re + a.im * a.im; }
static inline double cabs_(cx a) { return hypot(a.re, a.im); }
static inline double norm1(cx a) { return fabs(a.re) + fabs(a.im); }
static inline int is_zero(cx a) { return a.re == 0.0 && a.im == 0.0; }

/* 1/z: plain (c - id)/(c^2 + d^2) inside 1e-300 < c^2 + d^2 < 1e300 with d != 0, Smith's
 * algorithm otherwise (complex.rs:21-38). */
static inline cx inv(cx z) {
    double c = z.re, d = z.im;
    double s = c * c + d * d;
    if (s > 1e-300 && s < 1e300 && d != 0.0) {
        double t = 1.0 / s;
        return C(c * t, -(d * t));
    }
    if (fabs(d) <= fabs(c)) {
        double r = d / c;
        double t = 1.0 / (c + d * r);
        return C(t, -(r * t));
    } else {
        double r = c / d;
        double t = 1.0 / (c * r + d);
        return C(r * t, -t);
    }
}

/* e^w - 1 for Re w <= 0 (complex.rs:51-56). */
static inline cx exp_m1c(cx w) {
    double s = sin(w.im), c = cos(w.im);
    double h = sin(w.im * 0.5);
    return C(expm1(w.re) * c - h * h * 2.0, exp(w.re) * s);
}

/* num-complex sin/cos: sin(a+ib) = sin a cosh b + i cos a sinh b; cos(a+ib) = cos a cosh b -
 * i sin a sinh b. */
static inline cx csin_(cx z) { return C(sin(z.re) * cosh(z.im), cos(z.re) * sinh(z.im)); }
static inline cx ccos_(cx z) { return C(cos(z.re) * cosh(z.im), -(sin(z.re) * sinh(z.im))); }

/* ---- truncation and start index (mie/mod.rs:122-145, mie/kernel.rs:76-112) ---- */
int64_t orc_default_n_max(double x) {
    double n = ceil(x + 6.0 * cbrt(x) + 3.0);
    if (isfinite(n) && n < 9007199254740992.0) return (int64_t)n;
    return INT64_MAX;
}

static const double BITS = 52.0; /* -log2(f64::EPSILON) */

int64_t orc_start_index(int64_t n_top, double rho_abs) {
    int64_t k0 = n_top;
    int64_t c15 = (int64_t)ceil(1.5 * rho_abs);
    if (c15 > k0) k0 = c15;
    double q = (double)(2 * k0 + 1) / rho_abs - 0.5;
    int64_t enclosed = k0 + (int64_t)ceil(BITS / (2.0 * log2(q))) + 2;
    if ((double)n_top >= 1.5 * rho_abs) return enclosed;
    double s0 = (double)n_top - rho_abs;
    double s0_32 = s0 > 0.0 ? s0 * sqrt(s0) : 0.0;
    double s = pow(s0_32 + 1.4 * 0.37 * BITS * sqrt(rho_abs), 2.0 / 3.0);
    int64_t margin = (int64_t)ceil(BITS / 2.0) + 4;
    int64_t asymptotic = (int64_t)ceil(rho_abs + s) + margin;
    return asymptotic < enclosed ? asymptotic : enclosed;
}

/* One step of u_k = 1/(c_k - u_{k+1}), c_k = (2k+1)/z (mie/kernel.rs:116-136). */
static inline cx step_complex(int64_t k, cx inv_z, cx u) {
    cx c = scl(inv_z, (double)(2 * k + 1));
    cx den = sub(c, u);
    if (is_zero(den)) den = R(EPS * (cabs_(c) + cabs_(u)));
    return inv(den);
}

static inline double step_real(int64_t k, double inv_rho, double u) {
    double c = inv_rho * (double)(2 * k + 1);
    double den = c - u;
    if (den == 0.0) den = EPS * (fabs(c) + fabs(u));
    return 1.0 / den;
}

/* u_k(m x) for k = 1..n_top+1 into out[k] (mie/kernel.rs:159-205): the real recurrence when
 * Im m = 0. */
static void ratios(cx *out, cx m, double x, int64_t n_top) {
    cx z = scl(m, x);
    for (int64_t k = 0; k <= n_top + 1; k++) out[k] = R(0.0);
    if (m.im == 0.0) {
        double inv_rho = 1.0 / z.re;
        int64_t start = orc_start_index(n_top + 1, fabs(z.re));
        double u = 0.0;
        for (int64_t k = start; k >= 1; k--) {
            u = step_real(k, inv_rho, u);
            if (k <= n_top + 1) out[k] = R(u);
        }
    } else {
        cx inv_z = inv(z);
        int64_t start = orc_start_index(n_top + 1, cabs_(z));
        cx u = R(0.0);
        for (int64_t k = start; k >= 1; k--) {
            u = step_complex(k, inv_z, u);
            if (k <= n_top + 1) out[k] = u;
        }
    }
}

static int64_t upward_orders(double x, int64_t n) {
    if (x >= 1.0) {
        int64_t f = (int64_t)floor(x);
        return f < n ? f : n;
    }
    return 0;
}

/* State::prepare (mie/kernel.rs:249-313): u_k(mx) for k = 1..n+1 and u_k(x) above the upward
 * orders. The reference interleaves the two independent chains; run one after the other they
 * give the same values. */
static void prepare(double x, cx m, int64_t n, cx *u_z, double *u_x) {
    int64_t n_s = upward_orders(x, n);
    int64_t top = n + 1;
    for (int64_t k = 0; k <= n + 1; k++) u_x[k] = 0.0;
    if (n_s < n) {
        int64_t start = orc_start_index(top, fabs(x));
        int64_t stop = n_s + 1 > 1 ? n_s + 1 : 1;
        double inv_x = 1.0 / x, ux = 0.0;
        for (int64_t k = start; k >= stop; k--) {
            ux = step_real(k, inv_x, ux);
            if (k <= top) u_x[k] = ux;
        }
    }
    ratios(u_z, m, x, n);
}

/* One coefficient's effective log-derivative at the outer surface (mie/kernel.rs:16-38). */
typedef struct {
    int direct_a, direct_b; /* Side::Direct (else Side::Delta) */
    cx va, vb;
} outer_side;

typedef struct {
    cx hp[2], hplus[2], extra_m[2];
} outer_derivs;

typedef struct {
    cx *a, *b, *da_dx, *da_dm, *db_dx, *db_dm;
    double abs_sum, abs_jac[3];
    cx *stored; /* 6 per order, for layered derivatives */
} finish_out;

static inline double term_(cx f, cx df, cx l, double n2) {
    return -((df.im - 2.0 * f.im * l.re) * n2);
}

/* State::finish (mie/kernel.rs:325-559). */
static void finish(double x, cx m, int64_t n, const cx *u_z, const double *u_x, int derivs,
                   const outer_side *outer, const outer_derivs *od, finish_out *o) {
    const cx one = R(1.0);
    cx z = scl(m, x);
    int64_t n_s = upward_orders(x, n);
    double inv_x = 1.0 / x;
    cx inv_z = inv(z);
    cx inv_m = inv(m);
    cx m2 = mul(m, m);
    cx inv_m2 = mul(inv_m, inv_m);
    cx m2_minus_1 = mul(sub(m, one), add(m, one));
    cx c_a = neg(mul(m2_minus_1, inv_m2));
    double sin_x = sin(x), cos_x = cos(x);
    double psi_prev = cos_x, psi = sin_x, chi_prev = -sin_x, chi = cos_x;
    cx t = R(0.0), inv_xi = R(0.0), r = R(0.0);
    if (n_s == 0) {
        cx xi0 = C(sin_x, -cos_x);
        t = mul(R(sin_x), inv(xi0));
        inv_xi = C(sin_x, cos_x);
        r = C(0.0, 1.0);
    }
    int m_is_one = (m.re == 1.0 && m.im == 0.0);
    double abs_sum = 0.0, jac[3] = {0.0, 0.0, 0.0};
    for (int64_t k = 1; k <= n; k++) {
        double kf = (double)k;
        double k_x = kf * inv_x;
        cx d = sub(scl(inv_z, (double)(k + 1)), u_z[k + 1]);
        const outer_side *side = outer ? &outer[k - 1] : NULL;
        cx hh_a = d, hh_b = d;
        if (side) {
            hh_a = side->direct_a ? side->va : add(d, side->va);
            hh_b = side->direct_b ? side->vb : add(d, side->vb);
        }
        cx d_m = mul(hh_a, inv_m);
        cx md = mul(hh_b, m);
        cx big_a = add(d_m, R(k_x));
        cx big_b = add(md, R(k_x));
        cx an, bn, inv_da, inv_db, g_a, g_b, h_a, h_b;
        if (k <= n_s) {
            double c = (double)(2 * k - 1) * inv_x;
            double psi_next = c * psi - psi_prev;
            psi_prev = psi;
            psi = psi_next;
            double chi_next = c * chi - chi_prev;
            chi_prev = chi;
            chi = chi_next;
            cx xi = C(psi, -chi);
            cx xi_prev = C(psi_prev, -chi_prev);
            inv_da = inv(sub(mul(big_a, xi), xi_prev));
            inv_db = inv(sub(mul(big_b, xi), xi_prev));
            if (!side) {
                an = mul(sub(scl(big_a, psi), R(psi_prev)), inv_da);
                bn = mul(sub(scl(big_b, psi), R(psi_prev)), inv_db);
            } else {
                if (!side->direct_a) {
                    cx own = m_is_one ? R(0.0)
                                      : sub(scl(add(mul(d, inv_m), R(k_x)), psi), R(psi_prev));
                    an = mul(add(own, scl(mul(side->va, inv_m), psi)), inv_da);
                } else {
                    an = mul(sub(scl(big_a, psi), R(psi_prev)), inv_da);
                }
                if (!side->direct_b) {
                    cx own = m_is_one ? R(0.0)
                                      : sub(scl(add(mul(d, m), R(k_x)), psi), R(psi_prev));
                    bn = mul(add(own, scl(mul(m, side->vb), psi)), inv_db);
                } else {
                    bn = mul(sub(scl(big_b, psi), R(psi_prev)), inv_db);
                }
            }
            g_a = mul(xi, inv_da);
            g_b = mul(xi, inv_db);
            h_a = mul(xi_prev, inv_da);
            h_b = mul(xi_prev, inv_db);
            if (k == n_s) {
                inv_xi = inv(xi);
                t = mul(R(psi), inv_xi);
                r = mul(xi_prev, inv_xi);
            }
        } else {
            r = inv(sub(R((double)(2 * k - 1) * inv_x), r));
            t = scl(mul(t, r), u_x[k]);
            inv_xi = mul(inv_xi, r);
            cx u_z1 = u_z[k + 1];
            cx u_x1 = R(u_x[k + 1]);
            cx num_a = sub(add(scl(c_a, (double)(k + 1) * inv_x), u_x1), mul(u_z1, inv_m));
            cx num_b = sub(u_x1, mul(m, u_z1));
            if (side) {
                cx d_x = sub(R((double)(k + 1) * inv_x), u_x1);
                num_a = side->direct_a ? sub(d_m, d_x) : add(num_a, mul(side->va, inv_m));
                num_b = side->direct_b ? sub(md, d_x) : add(num_b, mul(m, side->vb));
            }
            cx ia = inv(sub(big_a, r));
            cx ib = inv(sub(big_b, r));
            an = mul(mul(t, num_a), ia);
            bn = mul(mul(t, num_b), ib);
            inv_da = mul(inv_xi, ia);
            inv_db = mul(inv_xi, ib);
            g_a = ia;
            g_b = ib;
            h_a = mul(r, ia);
            h_b = mul(r, ib);
        }
        o->a[k - 1] = an;
        o->b[k - 1] = bn;
        double na = norm2(inv_da), nb = norm2(inv_db);
        double weight = (double)(2 * k + 1);
        abs_sum += weight * (-(d_m.im * na) - md.im * nb);
        if (derivs && !od) {
            cx u = u_z[k + 1];
            cx one_u2 = add(one, mul(u, u));
            cx dp = sub(scl(sub(mul(scl(u, 2.0), inv_z), mul(inv_z, inv_z)), (double)(k + 1)),
                        one_u2);
            cx d_plus = sub(scl(u, weight), mul(z, one_u2));
            cx d_minus = sub(scl(d, 2.0), d_plus);
            double kk1 = kf * (double)(k + 1);
            cx ida2 = mul(inv_da, inv_da), idb2 = mul(inv_db, inv_db);
            o->da_dm[k - 1] = mul(mul(times_i(d_minus), inv_m2), ida2);
            o->db_dm[k - 1] = neg(mul(times_i(d_plus), idb2));
            o->da_dx[k - 1] = mul(mul(mul(times_i(m2_minus_1),
                                          add(mul(d, d), R(kk1 * inv_x * inv_x))),
                                      inv_m2),
                                  ida2);
            o->db_dx[k - 1] = mul(times_i(m2_minus_1), idb2);
            cx fa_m = neg(mul(d_minus, inv_m2));
            cx fb_m = d_plus;
            double k_x2 = k_x * inv_x;
            cx la_x = add(mul(g_a, add(sub(sub(dp, R(k_x2)), scl(big_a, k_x)), one)), mul(h_a, d_m));
            cx lb_x = add(mul(g_b, add(sub(sub(mul(m2, dp), R(k_x2)), scl(big_b, k_x)), one)),
                          mul(h_b, md));
            cx la_m = mul(fa_m, g_a), lb_m = mul(fb_m, g_b);
            jac[0] += weight * (term_(d_m, dp, la_x, na) + term_(md, mul(m2, dp), lb_x, nb));
            jac[1] += weight * (term_(d_m, fa_m, la_m, na) + term_(md, fb_m, lb_m, nb));
            jac[2] += weight * (term_(d_m, times_i(fa_m), times_i(la_m), na) +
                                term_(md, times_i(fb_m), times_i(lb_m), nb));
        }
        if (derivs && od) {
            const outer_derivs *q = &od[k - 1];
            double kk1 = kf * (double)(k + 1);
            cx ida2 = mul(inv_da, inv_da), idb2 = mul(inv_db, inv_db);
            cx dpa = q->hp[0], dpb = q->hp[1];
            cx minus_a = sub(scl(hh_a, 2.0), q->hplus[0]);
            cx fa_m = add(neg(mul(minus_a, inv_m2)), mul(q->extra_m[0], inv_m));
            cx fb_m = add(q->hplus[1], mul(m, q->extra_m[1]));
            o->da_dm[k - 1] = neg(mul(times_i(fa_m), ida2));
            o->db_dm[k - 1] = neg(mul(times_i(fb_m), idb2));
            o->da_dx[k - 1] = mul(mul(mul(times_i(m2_minus_1),
                                          add(mul(hh_a, hh_a), R(kk1 * inv_x * inv_x))),
                                      inv_m2),
                                  ida2);
            o->db_dx[k - 1] = mul(times_i(m2_minus_1), idb2);
            double k_x2 = k_x * inv_x;
            cx la_x = add(mul(g_a, add(sub(sub(dpa, R(k_x2)), scl(big_a, k_x)), one)), mul(h_a, d_m));
            cx lb_x = add(mul(g_b, add(sub(sub(mul(m2, dpb), R(k_x2)), scl(big_b, k_x)), one)),
                          mul(h_b, md));
            cx la_m = mul(fa_m, g_a), lb_m = mul(fb_m, g_b);
            jac[0] += weight * (term_(d_m, dpa, la_x, na) + term_(md, mul(m2, dpb), lb_x, nb));
            jac[1] += weight * (term_(d_m, fa_m, la_m, na) + term_(md, fb_m, lb_m, nb));
            jac[2] += weight * (term_(d_m, times_i(fa_m), times_i(la_m), na) +
                                term_(md, times_i(fb_m), times_i(lb_m), nb));
            cx *st = &o->stored[6 * (k - 1)];
            st[0] = inv_da; st[1] = inv_db; st[2] = g_a; st[3] = g_b; st[4] = d_m; st[5] = md;
        }
    }
    o->abs_sum = abs_sum;
    o->abs_jac[0] = jac[0];
    o->abs_jac[1] = jac[1];
    o->abs_jac[2] = jac[2];
}

int orc_mie_solve(double x, orc_cx m, int64_t n, int derivs, orc_cx *a, orc_cx *b,
                  orc_cx *da_dx, orc_cx *da_dm, orc_cx *db_dx, orc_cx *db_dm,
                  double *abs_sum, double *abs_jac) {
    cx *u_z = (cx *)malloc(sizeof(cx) * (size_t)(n + 2));
    double *u_x = (double *)malloc(sizeof(double) * (size_t)(n + 2));
    prepare(x, m, n, u_z, u_x);
    finish_out o = {a, b, da_dx, da_dm, db_dx, db_dm, 0.0, {0, 0, 0}, NULL};
    finish(x, m, n, u_z, u_x, derivs, NULL, NULL, &o);
    *abs_sum = o.abs_sum;
    if (abs_jac) memcpy(abs_jac, o.abs_jac, sizeof(o.abs_jac));
    free(u_z);
    free(u_x);
    return 0;
}

/* ---- efficiencies (mie/efficiencies.rs:47-230) ---- */
static inline double re_mul_conj(cx u, cx v) { return u.re * v.re + u.im * v.im; }

typedef struct {
    double sca, g;
    cx back;
} sums_t;

static sums_t value_sums(int64_t n, const cx *a, const cx *b) {
    sums_t s = {0.0, 0.0, R(0.0)};
    for (int64_t k = 1; k <= n; k++) {
        cx ak = a[k - 1], bk = b[k - 1];
        double w = (double)(2 * k + 1);
        s.sca += w * (norm2(ak) + norm2(bk));
        cx t = scl(sub(ak, bk), w);
        s.back = (k % 2 == 0) ? add(s.back, t) : sub(s.back, t);
        double kf = (double)k;
        if (k < n) {
            cx a1 = a[k], b1 = b[k];
            s.g += kf * (kf + 2.0) / (kf + 1.0) * (re_mul_conj(ak, a1) + re_mul_conj(bk, b1));
        }
        s.g += w / (kf * (kf + 1.0)) * re_mul_conj(ak, bk);
    }
    return s;
}

void orc_efficiencies(double x, int64_t n, const orc_cx *a, const orc_cx *b, double abs_sum,
                      double out[5]) {
    sums_t s = value_sums(n, a, b);
    double inv_x = 1.0 / x, inv_x2 = inv_x * inv_x, two_x2 = 2.0 * inv_x2;
    double q_sca = two_x2 * s.sca, q_abs = two_x2 * abs_sum;
    double g = s.sca == 0.0 ? 0.0 : 2.0 * s.g / s.sca;
    out[0] = q_sca + q_abs;
    out[1] = q_sca;
    out[2] = q_abs;
    out[3] = norm2(s.back) * inv_x2;
    out[4] = g;
}

void orc_efficiencies_jacobian(double x, int64_t n, const orc_cx *a, const orc_cx *b,
                               const orc_cx *da_dx, const orc_cx *da_dm, const orc_cx *db_dx,
                               const orc_cx *db_dm, double abs_sum, const double abs_jac[3],
                               int outer, double out[15]) {
    double inv_x = 1.0 / x, inv_x2 = inv_x * inv_x, two_x2 = 2.0 * inv_x2;
    sums_t values = value_sums(n, a, b);
    double q[5];
    orc_efficiencies(x, n, a, b, abs_sum, q);
    double sca[3] = {0, 0, 0}, g[3] = {0, 0, 0};
    cx back[3] = {R(0), R(0), R(0)};
    for (int64_t k = 1; k <= n; k++) {
        int64_t i = k - 1;
        cx ak = a[i], bk = b[i];
        cx dak[3] = {da_dx[i], da_dm[i], times_i(da_dm[i])};
        cx dbk[3] = {db_dx[i], db_dm[i], times_i(db_dm[i])};
        double w = (double)(2 * k + 1), kf = (double)k;
        double pair = kf * (kf + 2.0) / (kf + 1.0);
        double own = w / (kf * (kf + 1.0));
        for (int p = 0; p < 3; p++) {
            sca[p] += w * 2.0 * (re_mul_conj(dak[p], ak) + re_mul_conj(dbk[p], bk));
            cx t = scl(sub(dak[p], dbk[p]), w);
            back[p] = (k % 2 == 0) ? add(back[p], t) : sub(back[p], t);
            if (k < n) {
                cx a1 = a[k], b1 = b[k];
                cx da1[3] = {da_dx[k], da_dm[k], times_i(da_dm[k])};
                cx db1[3] = {db_dx[k], db_dm[k], times_i(db_dm[k])};
                g[p] += pair * (re_mul_conj(dak[p], a1) + re_mul_conj(ak, da1[p]) +
                                re_mul_conj(dbk[p], b1) + re_mul_conj(bk, db1[p]));
            }
            g[p] += own * (re_mul_conj(dak[p], bk) + re_mul_conj(ak, dbk[p]));
        }
    }
    for (int p = 0; p < 3; p++) {
        double xt = (p == 0 && outer) ? 2.0 * inv_x : 0.0;
        double q_sca = two_x2 * sca[p] - xt * q[1];
        double q_abs = two_x2 * abs_jac[p] - xt * q[2];
        double q_back = 2.0 * re_mul_conj(back[p], values.back) * inv_x2 - xt * q[3];
        double gg = values.sca == 0.0 ? 0.0 : (2.0 * g[p] - q[4] * sca[p]) / values.sca;
        out[p * 5 + 0] = q_sca + q_abs;
        out[p * 5 + 1] = q_sca;
        out[p * 5 + 2] = q_abs;
        out[p * 5 + 3] = q_back;
        out[p * 5 + 4] = gg;
    }
}

int orc_mie_efficiencies_batch(const double *x, const double *m, int64_t len, int64_t n_max,
                               double *values, double *jac, int threads) {
    int bad = 0;
    int64_t first = len;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads) reduction(| : bad)
#endif
    for (int64_t i = 0; i < len; i++) {
        double xi = x[i];
        cx mi = C(m[2 * i], m[2 * i + 1]);
        if (!(xi >= 1e-30 && xi <= 1e6) || !(mi.re >= 0 && mi.im >= 0)) {
            bad = 1;
#ifdef _OPENMP
#pragma omp critical
#endif
            if (i < first) first = i;
            continue;
        }
        int64_t n = n_max > 0 ? n_max : orc_default_n_max(xi);
        int derivs = jac != NULL;
        cx *buf = (cx *)malloc(sizeof(cx) * (size_t)n * (derivs ? 6 : 2));
        cx *a = buf, *b = buf + n;
        cx *dx_a = derivs ? buf + 2 * n : NULL, *dm_a = derivs ? buf + 3 * n : NULL;
        cx *dx_b = derivs ? buf + 4 * n : NULL, *dm_b = derivs ? buf + 5 * n : NULL;
        double abs_sum, abs_jac[3];
        orc_mie_solve(xi, mi, n, derivs, a, b, dx_a, dm_a, dx_b, dm_b, &abs_sum, abs_jac);
        double q[5];
        orc_efficiencies(xi, n, a, b, abs_sum, q);
        for (int k = 0; k < 5; k++) values[k * len + i] = q[k];
        if (derivs) {
            double j[15];
            orc_efficiencies_jacobian(xi, n, a, b, dx_a, dm_a, dx_b, dm_b, abs_sum, abs_jac, 1, j);
            for (int k = 0; k < 15; k++) jac[k * len + i] = j[k];
        }
        free(buf);
    }
    (void)threads;
    return bad ? -(int)(first + 1) : 0;
}

/* ---- amplitudes (mie/amplitudes.rs:52-65 and the sums at :300-320) ---- */
void orc_amplitudes(int64_t n, const orc_cx *a, const orc_cx *b, const double *mu, int64_t n_mu,
                    orc_cx *s1, orc_cx *s2) {
    for (int64_t j = 0; j < n_mu; j++) {
        double muj = mu[j];
        cx t1 = R(0.0), t2 = R(0.0);
        double pi_prev = 0.0, pi = 1.0;
        for (int64_t k = 1; k <= n; k++) {
            double nf = (double)k, n1 = nf + 1.0;
            double tau = nf * muj * pi - n1 * pi_prev;
            double w = (double)(2 * k + 1) / (nf * n1);
            double wp = w * pi, wt = w * tau;
            cx an = a[k - 1], bn = b[k - 1];
            t1 = add(add(t1, scl(an, wp)), scl(bn, wt));
            t2 = add(add(t2, scl(an, wt)), scl(bn, wp));
            double pi_next = ((double)(2 * k + 1) * muj * pi - n1 * pi_prev) / nf;
            pi_prev = pi;
            pi = pi_next;
        }
        s1[j] = t1;
        s2[j] = t2;
    }
}

/* ---- Gauss-Legendre (special/gauss_legendre.rs:210-216) ----
 * The reference computes the rule by Newton in double-double up to 100 points and by
 * Bogaert's asymptotic expansions above; this restatement runs Newton on the three-term
 * recurrence in long double (64-bit significand), which rounds to within an ulp or so of the
 * f64 rule, and mirrors the positive half so the rule is exactly symmetric. */
void orc_gauss_legendre(int64_t n, double *nodes, double *weights) {
    const long double pi_l = 3.14159265358979323846264338327950288L;
    for (int64_t i = 0; i < (n + 1) / 2; i++) {
        /* i-th largest root, Tricomi's start. */
        long double th = pi_l * ((long double)i + 0.75L) / ((long double)n + 0.5L);
        long double x = cosl(th);
        long double dp = 0.0L;
        for (int it = 0; it < 100; it++) {
            long double p0 = 1.0L, p1 = x;
            for (int64_t k = 2; k <= n; k++) {
                long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / (long double)k;
                p0 = p1;
                p1 = p2;
            }
            if (n == 1) { p1 = x; p0 = 1.0L; }
            dp = (long double)n * (x * p1 - p0) / (x * x - 1.0L);
            long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) <= 1e-19L * fabsl(x) || fabsl(dx) < 1e-21L) {
                /* one more evaluation of the derivative at the converged root */
                p0 = 1.0L; p1 = x;
                for (int64_t k = 2; k <= n; k++) {
                    long double p2 = ((2.0L * k - 1.0L) * x * p1 - (k - 1.0L) * p0) / (long double)k;
                    p0 = p1;
                    p1 = p2;
                }
                if (n == 1) { p1 = x; p0 = 1.0L; }
                dp = (long double)n * (x * p1 - p0) / (x * x - 1.0L);
                break;
            }
        }
        if (n % 2 == 1 && i == (n - 1) / 2) x = 0.0L;
        long double w = 2.0L / ((1.0L - x * x) * dp * dp);
        if (n == 1) w = 2.0L;
        /* ascending: the i-th largest positive root sits at n-1-i, its mirror at i */
        nodes[n - 1 - i] = (double)x;
        nodes[i] = -(double)x;
        weights[n - 1 - i] = (double)w;
        weights[i] = (double)w;
    }
}

/* ---- Wigner d (special/wigner.rs:152-460), for small |m|, |n| (no scaled start) ---- */
static void start_value(int64_t a, int64_t b, double x, double *v_out) {
    double half_sin = sqrt((1.0 - x) * 0.5);
    double half_cos = sqrt((1.0 + x) * 0.5);
    if (a >= 2 && half_sin == 0.0) {
        if (a > 2) { *v_out = 0.0; return; }
        double b1 = (double)b + 1.0, b2 = (double)b + 2.0;
        double v = (1.0 - x) * 0.5 * sqrt(b1 * b2 * 0.5);
        for (int64_t k = 0; k < b; k++) v *= half_cos;
        *v_out = v;
        return;
    }
    double v = 1.0;
    int64_t top = a > b ? a : b;
    for (int64_t k = 1; k <= top; k++) {
        if (k <= a) v = v * sqrt((double)(b + k) / (double)k) * half_sin;
        if (k <= b) v *= half_cos;
    }
    *v_out = v;
}

static void three_term(int64_t m, int64_t n, double x, int64_t l_min, int64_t len, double *out) {
    int64_t a = llabs(m - n), b = llabs(m + n);
    double mag;
    start_value(a, b, x, &mag);
    double start = (n >= m || (m - n) % 2 == 0) ? mag : -mag;
    double mf = (double)m, nf = (double)n, q = 0.0;
    double prev = 0.0, cur = start;
    int64_t s = l_min;
    out[s] = cur;
    while (s + 1 < len) {
        double sf = (double)s;
        double c_prev = q * (sf + 1.0);
        double numer = (x * (sf * (sf + 1.0)) - mf * nf) * (2.0 * sf + 1.0) * cur - c_prev * prev;
        double s1 = sf + 1.0;
        q = sqrt((s1 - mf) * (s1 + mf)) * sqrt((s1 - nf) * (s1 + nf));
        double next = numer / (q * sf);
        prev = cur;
        cur = next;
        s++;
        out[s] = cur;
    }
}

static void difference_run(double mf, double w, int64_t s, double delta, double cur,
                           int64_t len, double *out) {
    out[s] = cur;
    while (s + 1 < len) {
        double sf = (double)s, s1 = sf + 1.0;
        double k = (2.0 * sf + 1.0) * sf * s1;
        double b = s1 * ((sf - mf) * (sf + mf));
        double c = sf * ((s1 - mf) * (s1 + mf));
        double next = (w * k * cur + delta * b) * (1.0 / c);
        delta = next;
        cur = cur + next;
        s++;
        out[s] = cur;
    }
}

static void difference_form(int64_t m, double x, int64_t l_min, int64_t len, double *out) {
    int64_t b = 2 * llabs(m);
    double start;
    start_value(0, b, x, &start);
    double w = x - 1.0;
    if (l_min == 0) {
        out[0] = start;
        if (len > 1) difference_run((double)m, w, 1, w, x, len, out);
    } else {
        difference_run((double)m, w, l_min, start, start, len, out);
    }
}

static void general_difference_form(int64_t m, int64_t n, double x, int64_t l_min, int64_t len,
                                    double *out) {
    int64_t a = llabs(m - n), b = llabs(m + n);
    double mag;
    start_value(a, b, x, &mag);
    double start = (n >= m || (m - n) % 2 == 0) ? mag : -mag;
    double mf = (double)m, nf = (double)n, lf = (double)l_min;
    double w = x - 1.0, q = 0.0, u = lf * (lf + 1.0) - mf * nf;
    double delta = start, cur = start;
    int64_t s = l_min;
    out[s] = cur;
    while (s + 1 < len) {
        double sf = (double)s, s1 = sf + 1.0;
        double q1 = sqrt(((s1 - mf) * (s1 + mf)) * ((s1 - nf) * (s1 + nf)));
        double alpha = sf * s1 - mf * nf;
        double dmn = mf - nf;
        double numer = (s1 * s1) * (1.0 + dmn * dmn - 2.0 * s1) + (2.0 * s1) * (mf * nf);
        double v = numer / (alpha + q1);
        double k = w * ((2.0 * sf + 1.0) * sf * s1) + u * s1 + v * sf;
        double next = (k * cur + delta * (q * s1)) * (1.0 / (q1 * sf));
        q = q1;
        u = v + 2.0 * s1;
        delta = next;
        cur = cur + next;
        s++;
        out[s] = cur;
    }
}

int orc_wigner_d(int m_in, int n_in, double x, int64_t lmax, double *out) {
    int64_t m = m_in, n = n_in;
    int ok = (m == 0 && n == 0) || (m == 2 && n == 2) || (m == 2 && n == -2) || (m == 0 && n == 2);
    if (!ok) return -1;
    int64_t len = lmax + 1;
    int64_t l_min = llabs(m) > llabs(n) ? llabs(m) : llabs(n);
    if (l_min >= len) {
        for (int64_t l = 0; l < len; l++) out[l] = 0.0;
        return 0;
    }
    for (int64_t l = 0; l < l_min; l++) out[l] = 0.0;
    int reflect = x < 0.0 && (m == -n || -x >= 0.9);
    if (reflect) { n = -n; x = -x; }
    if (m == n && x >= 0.0) difference_form(m, x, l_min, len, out);
    else if (x >= 0.9) general_difference_form(m, n, x, l_min, len, out);
    else three_term(m, n, x, l_min, len, out);
    if (reflect) {
        int64_t r = (l_min + m) % 2;
        if (r < 0) r += 2;
        int64_t first_odd = 1 - r;
        for (int64_t i = first_odd; l_min + i < len; i += 2) out[l_min + i] = -out[l_min + i];
    }
    return 0;
}

/* ---- Greek projection (mie/greek.rs:161-221) ---- */
void orc_greek_project(int64_t n_q, const double *nodes, const double *weights,
                       const double *f11, const double *f12, const double *f33,
                       const double *f34, int64_t order, double *out) {
    int64_t len = order + 1;
    double *a1 = out, *a2 = out + len, *a3 = out + 2 * len, *a4 = out + 3 * len;
    double *b1 = out + 4 * len, *b2 = out + 5 * len;
    for (int64_t l = 0; l < 6 * len; l++) out[l] = 0.0;
    double *d = (double *)malloc(sizeof(double) * 4 * (size_t)len);
    double *d00 = d, *d22 = d + len, *d2m2 = d + 2 * len, *d02 = d + 3 * len;
    for (int64_t k = 0; k < n_q; k++) {
        double mu = nodes[k], w = weights[k];
        orc_wigner_d(0, 0, mu, order, d00);
        orc_wigner_d(2, 2, mu, order, d22);
        orc_wigner_d(2, -2, mu, order, d2m2);
        orc_wigner_d(0, 2, mu, order, d02);
        double F11 = w * f11[k], F33 = w * f33[k];
        double sum = F11 + F33, diff = F11 - F33;
        double F12 = w * f12[k], F34 = w * f34[k];
        for (int64_t l = 0; l < len; l++) {
            a1[l] += F11 * d00[l];
            a4[l] += F33 * d00[l];
            a2[l] += sum * d22[l];
            a3[l] += diff * d2m2[l];
            b1[l] += F12 * d02[l];
            b2[l] += F34 * d02[l];
        }
    }
    for (int64_t l = 0; l < len; l++) {
        double s = (double)(2 * l + 1) * 0.5;
        double plus = a2[l], minus = a3[l];
        a1[l] *= s;
        a4[l] *= s;
        a2[l] = s * 0.5 * (plus + minus);
        a3[l] = s * 0.5 * (plus - minus);
        b1[l] *= s;
        b2[l] *= s;
    }
    free(d);
}

/* ---- layered spheres (layered/kernel.rs:28-587) ---- */
typedef struct {
    cx z, inv_z, pi, w, w_n;
} arg_t;

static inline cx two_i(cx z) { return C(-(z.im * 2.0), z.re * 2.0); }

static arg_t arg_new(cx z, int chi) {
    arg_t a;
    a.z = z;
    a.inv_z = inv(z);
    if (chi) {
        cx s = csin_(z), c = ccos_(z);
        a.pi = mul(s, c);
        a.w = add(a.inv_z, mul(s, inv(c)));
    } else {
        a.pi = neg(scl(exp_m1c(two_i(z)), 0.5));
        a.w = sub(a.inv_z, C(0.0, 1.0));
    }
    a.w_n = R(0.0);
    return a;
}

static inline void arg_step(arg_t *a, int64_t n, cx u_n) {
    a->w_n = a->w;
    a->pi = mul(mul(a->pi, u_n), a->w);
    a->w = sub(scl(a->inv_z, (double)(2 * n + 1)), inv(a->w));
}
static inline cx arg_d1(const arg_t *a, int64_t n, cx u_next) {
    return sub(scl(a->inv_z, (double)(n + 1)), u_next);
}
static inline cx arg_d3(const arg_t *a, int64_t n) { return sub(scl(a->inv_z, (double)(n + 1)), a->w); }
static inline cx arg_d1_prime(const arg_t *a, int64_t n, cx u) {
    cx iz = a->inv_z;
    return sub(scl(sub(mul(scl(u, 2.0), iz), mul(iz, iz)), (double)(n + 1)), add(R(1.0), mul(u, u)));
}
static inline cx arg_d1_plus(const arg_t *a, int64_t n, cx u) {
    return sub(scl(u, (double)(2 * n + 1)), mul(a->z, add(R(1.0), mul(u, u))));
}

typedef struct {
    cx u_prev, u1, d1i, d3i, d1o, d3o, pio, p, w;
    int derivatives;
    int64_t n;
    cx u2, s;
    arg_t a1, a2;
} lstep_t;

typedef struct {
    cx hv, r2, em, hp, hplus;
} sens_t;

/* layered/kernel.rs:158-232. Returns h_out; *direct and *val describe the Side. */
static cx side_fn(cx mi, cx mo, int prev_direct, cx prev_val, cx h, cx lead, const lstep_t *s,
                  int lead_is_a, int *direct, cx *val, sens_t *sens) {
    (void)prev_val;
    cx mh = mul(mi, h);
    cx g1 = prev_direct ? sub(mh, mul(mo, s->d1i))
                        : add(add(sub(lead, mul(mi, s->u_prev)), mul(mo, s->u1)), mul(mi, prev_val));
    cx g2 = sub(mh, mul(mo, s->d3i));
    cx inv_den = inv(sub(mul(s->pio, g2), mul(s->p, g1)));
    cx delta = neg(mul(mul(mul(mul(s->w, s->p), g1), inv_den), inv(s->pio)));
    cx hd = add(s->d1o, delta);
    cx h_out;
    double lim = norm1(s->d1o) > norm1(delta) ? norm1(s->d1o) : norm1(delta);
    if (norm1(hd) * 4.0 >= lim) {
        *direct = 0;
        *val = delta;
        h_out = hd;
    } else {
        cx hp = sub(s->d3o, mul(mul(s->w, g2), inv_den));
        *direct = 1;
        *val = hp;
        h_out = hp;
    }
    if (!s->derivatives) return h_out;
    cx inv_mo = inv(mo);
    cx t = mul(mh, inv_mo);
    cx r = neg(mul(mul(mul(mo, s->w), s->s), inv_den));
    cx e = mul(g1, inv_mo);
    double sigma = lead_is_a ? 1.0 : -1.0;
    cx d1 = s->d1i;
    cx z1 = s->a1.z;
    cx own = lead_is_a ? sub(scl(d1, 2.0), arg_d1_plus(&s->a1, s->n, s->u1))
                       : neg(arg_d1_plus(&s->a1, s->n, s->u1));
    cx em;
    double lim2 = norm1(d1) > norm1(e) ? norm1(d1) : norm1(e);
    if (norm1(t) * 4.0 >= lim2) {
        em = add(own, mul(e, add(R(sigma), mul(z1, add(scl(d1, 2.0), e)))));
    } else {
        cx iz = s->a1.inv_z;
        cx v = sub(mul(mul(R((double)s->n * (double)(s->n + 1)), iz), iz), R(1.0));
        em = sub(scl(t, sigma), mul(z1, sub(v, mul(t, t))));
    }
    cx hp, hplus;
    if (!*direct) {
        cx d2 = add(scl(s->d1o, 2.0), delta);
        hp = sub(arg_d1_prime(&s->a2, s->n, s->u2), mul(delta, d2));
        hplus = add(arg_d1_plus(&s->a2, s->n, s->u2), mul(delta, sub(R(1.0), mul(s->a2.z, d2))));
    } else {
        cx hv = *val;
        cx iz = s->a2.inv_z;
        cx v = sub(mul(mul(R((double)s->n * (double)(s->n + 1)), iz), iz), R(1.0));
        hp = sub(v, mul(hv, hv));
        hplus = add(hv, mul(s->a2.z, hp));
    }
    sens->hv = h_out;
    sens->r2 = mul(r, r);
    sens->em = em;
    sens->hp = hp;
    sens->hplus = hplus;
    return h_out;
}

int orc_layered_solve(int layers, const double *x, const orc_cx *m, int64_t n, int derivs,
                      orc_cx *a, orc_cx *b, orc_cx *dgrp, double *abs_sum, double *abs_jac) {
    int last = layers - 1;
    if (layers < 1 || layers > 64) return -1;
    if (layers == 1) {
        double aj[3];
        orc_mie_solve(x[0], m[0], n, derivs, a, b, derivs ? dgrp : NULL,
                      derivs ? dgrp + n : NULL, derivs ? dgrp + 2 * n : NULL,
                      derivs ? dgrp + 3 * n : NULL, abs_sum, aj);
        if (derivs && abs_jac) memcpy(abs_jac, aj, sizeof(aj));
        return 0;
    }
    size_t np2 = (size_t)(n + 2);
    cx *u_z = (cx *)malloc(sizeof(cx) * np2);
    double *u_x = (double *)malloc(sizeof(double) * np2);
    prepare(x[last], m[last], n, u_z, u_x);
    /* u[0]: core m0 x0; u[2l-1]: m_l x_{l-1}; u[2l] (l < last): m_l x_l */
    cx *u = (cx *)malloc(sizeof(cx) * np2 * (size_t)(2 * last));
    ratios(u, m[0], x[0], n);
    for (int l = 1; l < layers; l++) {
        ratios(u + np2 * (size_t)(2 * l - 1), m[l], x[l - 1], n);
        if (l < last) ratios(u + np2 * (size_t)(2 * l), m[l], x[l], n);
    }
#define U(idx) (u + np2 * (size_t)(idx))
    arg_t *inner = (arg_t *)malloc(sizeof(arg_t) * (size_t)last);
    arg_t *outer_args = (arg_t *)malloc(sizeof(arg_t) * (size_t)last);
    cx *p = (cx *)malloc(sizeof(cx) * (size_t)last);
    cx *lead = (cx *)malloc(sizeof(cx) * (size_t)last);
    int *chi = (int *)malloc(sizeof(int) * (size_t)last);
    cx *sv = (cx *)malloc(sizeof(cx) * (size_t)last);
    for (int l = 1; l < layers; l++) {
        int c = m[l].im * x[l] < 0.5;
        chi[l - 1] = c;
        cx aa = m[l], bb = m[l - 1];
        lead[l - 1] = mul(mul(sub(aa, bb), add(aa, bb)), inv(mul(aa, bb)));
        cx z1 = scl(m[l], x[l - 1]), z2 = scl(m[l], x[l]);
        inner[l - 1] = arg_new(z1, c);
        outer_args[l - 1] = arg_new(z2, c);
        cx p0;
        if (c) {
            cx c2 = ccos_(z2);
            p0 = mul(mul(mul(csin_(z1), inv(ccos_(z1))), c2), c2);
        } else {
            double dx = x[l] - x[l - 1];
            double ea = -(m[l].im * dx * 2.0), eb = m[l].re * dx * 2.0;
            cx e = scl(C(cos(eb), sin(eb)), exp(ea));
            p0 = neg(scl(mul(e, exp_m1c(two_i(z1))), 0.5));
        }
        p[l - 1] = p0;
        if (derivs) {
            cx s0;
            if (c) {
                s0 = mul(ccos_(z2), inv(ccos_(z1)));
            } else {
                double dx = x[l] - x[l - 1];
                double bb2 = m[l].re * dx;
                s0 = scl(C(cos(bb2), sin(bb2)), exp(-(m[l].im * dx)));
            }
            sv[l - 1] = s0;
        }
    }
    arg_t core = arg_new(scl(m[0], x[0]), 0);
    outer_side *outer = (outer_side *)malloc(sizeof(outer_side) * (size_t)n);
    outer_derivs *od = derivs ? (outer_derivs *)malloc(sizeof(outer_derivs) * (size_t)n) : NULL;
    sens_t *sens_a = (sens_t *)malloc(sizeof(sens_t) * (size_t)last);
    sens_t *sens_b = (sens_t *)malloc(sizeof(sens_t) * (size_t)last);
    /* dh[l][j][k]: j = dHa/dx_l, dHa/dm_l, dHb/dx_l, dHb/dm_l for inner layers l < last */
    cx *dh = derivs ? (cx *)malloc(sizeof(cx) * (size_t)last * 4 * (size_t)n) : NULL;
    for (int64_t k = 1; k <= n; k++) {
        for (int l = 1; l < layers; l++) {
            cx *u1 = U(2 * l - 1);
            cx *u2 = (l == last) ? u_z : U(2 * l);
            arg_t *ai = &inner[l - 1], *ao = &outer_args[l - 1];
            arg_step(ai, k, u1[k]);
            arg_step(ao, k, u2[k]);
            p[l - 1] = mul(mul(mul(mul(p[l - 1], u1[k]), inv(ai->w_n)), ao->w_n), ao->w_n);
            if (derivs) sv[l - 1] = mul(mul(sv[l - 1], ao->w_n), inv(ai->w_n));
        }
        cx ha = arg_d1(&core, k, U(0)[k + 1]);
        cx hb = ha;
        int da = 0, db = 0;
        cx va = R(0.0), vb = R(0.0);
        for (int l = 1; l < layers; l++) {
            arg_t *ai = &inner[l - 1], *ao = &outer_args[l - 1];
            cx *u_prev = (l == 1) ? U(0) : U(2 * (l - 1));
            cx *u2 = (l == last) ? u_z : U(2 * l);
            lstep_t s;
            s.u_prev = u_prev[k + 1];
            s.u1 = U(2 * l - 1)[k + 1];
            s.d1i = arg_d1(ai, k, U(2 * l - 1)[k + 1]);
            s.d3i = arg_d3(ai, k);
            s.d1o = arg_d1(ao, k, u2[k + 1]);
            s.d3o = arg_d3(ao, k);
            s.pio = ao->pi;
            s.p = p[l - 1];
            s.w = chi[l - 1] ? R(-1.0) : C(0.0, 1.0);
            s.derivatives = derivs;
            s.n = k;
            s.u2 = u2[k + 1];
            s.s = derivs ? sv[l - 1] : R(0.0);
            s.a1 = *ai;
            s.a2 = *ao;
            cx ld = scl(lead[l - 1], (double)(k + 1) / x[l - 1]);
            int nda, ndb;
            cx nva, nvb;
            cx a_h = side_fn(m[l], m[l - 1], da, va, ha, ld, &s, 1, &nda, &nva, &sens_a[l - 1]);
            cx b_h = side_fn(m[l - 1], m[l], db, vb, hb, R(0.0), &s, 0, &ndb, &nvb, &sens_b[l - 1]);
            ha = a_h;
            hb = b_h;
            da = nda; va = nva; db = ndb; vb = nvb;
        }
        outer[k - 1].direct_a = da;
        outer[k - 1].va = va;
        outer[k - 1].direct_b = db;
        outer[k - 1].vb = vb;
        if (derivs) {
            cx u0 = U(0)[k + 1];
            cx core_h = arg_d1(&core, k, u0), core_hplus = arg_d1_plus(&core, k, u0);
            /* adjoint sweep (layered/kernel.rs:479-543) */
            double nn1 = (double)k * (double)(k + 1);
            cx extra[2];
            for (int sd = 0; sd < 2; sd++) {
                cx dxv[64], dmv[64];
                for (int l = 0; l < layers; l++) { dxv[l] = R(0.0); dmv[l] = R(0.0); }
                cx lambda = R(1.0);
                for (int l = last; l >= 1; l--) {
                    sens_t e = sd == 0 ? sens_a[l - 1] : sens_b[l - 1];
                    cx ml = m[l], mp = m[l - 1];
                    dmv[l] = add(dmv[l], mul(mul(mul(lambda, e.r2), e.em), inv(ml)));
                    cx h, hplus;
                    if (l == 1) { h = core_h; hplus = core_hplus; }
                    else {
                        sens_t pp = sd == 0 ? sens_a[l - 2] : sens_b[l - 2];
                        h = pp.hv; hplus = pp.hplus;
                    }
                    cx c = mul(sub(ml, mp), add(ml, mp));
                    cx w = mul(lambda, e.r2);
                    cx inv_mp = inv(mp);
                    cx jump, kk;
                    if (sd == 0) {
                        double inv_xl = 1.0 / x[l - 1];
                        jump = mul(mul(ml, c), add(mul(R(nn1 * inv_xl * inv_xl), inv(mul(mul(mul(mp, mp), ml), ml))),
                                                   mul(mul(mul(h, h), inv_mp), inv_mp)));
                        kk = neg(mul(mul(mul(ml, inv_mp), inv_mp), sub(scl(h, 2.0), hplus)));
                    } else {
                        cx inv_ml = inv(ml);
                        jump = mul(c, inv_ml);
                        kk = mul(hplus, inv_ml);
                    }
                    dxv[l - 1] = add(dxv[l - 1], mul(w, jump));
                    dmv[l - 1] = add(dmv[l - 1], mul(w, kk));
                    cx mi = sd == 0 ? ml : mp, mo = sd == 0 ? mp : ml;
                    lambda = mul(mul(w, mi), inv(mo));
                }
                extra[sd] = dmv[last];
                for (int l = 0; l < last; l++) {
                    dh[((size_t)l * 4 + 2 * sd) * (size_t)n + (size_t)(k - 1)] = dxv[l];
                    dh[((size_t)l * 4 + 2 * sd + 1) * (size_t)n + (size_t)(k - 1)] = dmv[l];
                }
            }
            od[k - 1].hp[0] = sens_a[last - 1].hp;
            od[k - 1].hp[1] = sens_b[last - 1].hp;
            od[k - 1].hplus[0] = sens_a[last - 1].hplus;
            od[k - 1].hplus[1] = sens_b[last - 1].hplus;
            od[k - 1].extra_m[0] = extra[0];
            od[k - 1].extra_m[1] = extra[1];
        }
    }
    cx *stored = derivs ? (cx *)malloc(sizeof(cx) * 6 * (size_t)n) : NULL;
    cx *og = derivs ? dgrp + (size_t)last * 4 * (size_t)n : NULL;
    finish_out o = {a, b, og, og ? og + n : NULL, og ? og + 2 * n : NULL, og ? og + 3 * n : NULL,
                    0.0, {0, 0, 0}, stored};
    finish(x[last], m[last], n, u_z, u_x, derivs, outer, od, &o);
    *abs_sum = o.abs_sum;
    if (derivs) {
        abs_jac[last * 3 + 0] = o.abs_jac[0];
        abs_jac[last * 3 + 1] = o.abs_jac[1];
        abs_jac[last * 3 + 2] = o.abs_jac[2];
        /* inner_grads (layered/kernel.rs:549-587) */
        cx mL = m[last];
        cx inv_mL = inv(mL);
        for (int l = 0; l < last; l++) {
            double jac[3] = {0, 0, 0};
            cx *g = dgrp + (size_t)l * 4 * (size_t)n;
            const cx *dhl = dh + (size_t)l * 4 * (size_t)n;
            for (int64_t k = 0; k < n; k++) {
                cx *st = &stored[6 * k];
                cx inv_da = st[0], inv_db = st[1], g_a = st[2], g_b = st[3], b_a = st[4], b_b = st[5];
                cx ida2 = mul(inv_da, inv_da), idb2 = mul(inv_db, inv_db);
                double na = norm2(inv_da), nb = norm2(inv_db);
                cx ba_x = mul(dhl[0 * n + k], inv_mL);
                cx ba_m = mul(dhl[1 * n + k], inv_mL);
                cx bb_x = mul(mL, dhl[2 * n + k]);
                cx bb_m = mul(mL, dhl[3 * n + k]);
                g[0 * n + k] = neg(mul(times_i(ba_x), ida2));
                g[1 * n + k] = neg(mul(times_i(ba_m), ida2));
                g[2 * n + k] = neg(mul(times_i(bb_x), idb2));
                g[3 * n + k] = neg(mul(times_i(bb_m), idb2));
                double weight = (double)(2 * k + 3);
#define PAIR(dA, dB) (term_(b_a, dA, mul(dA, g_a), na) + term_(b_b, dB, mul(dB, g_b), nb))
                jac[0] += weight * PAIR(ba_x, bb_x);
                jac[1] += weight * PAIR(ba_m, bb_m);
                jac[2] += weight * PAIR(times_i(ba_m), times_i(bb_m));
#undef PAIR
            }
            abs_jac[l * 3 + 0] = jac[0];
            abs_jac[l * 3 + 1] = jac[1];
            abs_jac[l * 3 + 2] = jac[2];
        }
    }
    int lossless = 1;
    for (int l = 0; l < layers; l++) if (m[l].im != 0.0) lossless = 0;
    if (lossless) {
        *abs_sum = 0.0;
        if (derivs) {
            for (int l = 0; l < layers; l++) {
                abs_jac[l * 3 + 0] = 0.0;
                abs_jac[l * 3 + 1] = 0.0;
            }
        }
    }
#undef U
    free(u_z); free(u_x); free(u); free(inner); free(outer_args); free(p); free(lead); free(chi);
    free(sv); free(outer); free(od); free(sens_a); free(sens_b); free(dh); free(stored);
    return 0;
}

#define GAMMA_TAIL_MG 37.0
/* ---- size distributions (ensemble/distribution.rs:15-68,293-430) ----
 * Parameters p[3] per kind: lognormal (r_g, sigma_g, -), gamma (a = r_eff, b = v_eff, -),
 * modified gamma (alpha, b, gamma), power law (alpha; r_min, r_max fixed), monodisperse
 * (radius, -, -). */
static double dist_reference(int kind, const double *p) {
    if (kind == ORC_LOGNORMAL) return p[0];
    if (kind == ORC_GAMMA) {
        double alpha = (1.0 - 3.0 * p[1]) / p[1];
        return alpha > 0.0 ? alpha * p[0] * p[1] : p[0];
    }
    if (kind == ORC_MODIFIED_GAMMA) return pow(fmax(p[0], 1.0) / (p[1] * p[2]), 1.0 / p[2]);
    if (kind == ORC_POWER_LAW) return p[0] >= 0.0 ? p[1] : p[2];
    return 1.0;
}

double orc_density(int kind, const double *p, double r, double grad[3]) {
    double r0 = dist_reference(kind, p);
    double ln_u = 0.0, d[3] = {0.0, 0.0, 0.0};
    if (kind == ORC_LOGNORMAL) {
        double mu = log(p[0]), s = log(p[1]);
        double z = log(r) - mu;
        ln_u = -z * z / (2.0 * s * s) - log(s) - log(r / r0);
        d[0] = z / (s * s * p[0]);
        d[1] = (z * z / (s * s * s) - 1.0 / s) / p[1];
    } else if (kind == ORC_GAMMA) {
        double a = p[0], b = p[1];
        double alpha = (1.0 - 3.0 * b) / b;
        double scale = a * b;
        double l = log(r / r0);
        ln_u = alpha * l - (r - r0) / scale;
        d[0] = (r - r0) * b / (scale * scale);
        d[1] = -l / (b * b) + (r - r0) * a / (scale * scale);
    } else if (kind == ORC_MODIFIED_GAMMA) {
        double alpha = p[0], b = p[1], g = p[2];
        double rg = pow(r, g), r0g = pow(r0, g);
        double l = log(r / r0);
        ln_u = alpha * l - b * (rg - r0g);
        d[0] = l;
        d[1] = -(rg - r0g);
        d[2] = -b * (rg * log(r) - r0g * log(r0));
    } else if (kind == ORC_POWER_LAW) {
        double l = log(r / r0);
        ln_u = -p[0] * l;
        d[0] = -l;
    } else if (kind == ORC_SEGMENT) {
        double s = (r - p[0]) / (p[1] - p[0]);
        ln_u = log((1.0 - p[2]) * (1.0 - s) + p[2] * s);
    }
    double u = exp(ln_u);
    if (grad)
        for (int k = 0; k < 3; k++) grad[k] = u * d[k];
    return u;
}

/* Whether the parameters are valid (distribution.rs:99-185). */
static int dist_valid(int kind, const double *p) {
    if (kind == ORC_LOGNORMAL) return p[0] > 0.0 && isfinite(p[0]) && p[1] > 1.0 && isfinite(p[1]);
    if (kind == ORC_GAMMA) return p[0] > 0.0 && isfinite(p[0]) && p[1] > 0.0 && p[1] < 0.5;
    if (kind == ORC_MODIFIED_GAMMA)
        return isfinite(p[0]) && p[0] > -1.0 && p[1] > 0.0 && isfinite(p[1]) && p[2] > 0.0 &&
               isfinite(p[2]);
    if (kind == ORC_POWER_LAW)
        return isfinite(p[0]) && p[1] > 0.0 && isfinite(p[1]) && p[2] > p[1] && isfinite(p[2]);
    if (kind == ORC_MONODISPERSE) return p[0] > 0.0 && isfinite(p[0]);
    if (kind == ORC_SEGMENT)
        return p[0] > 0.0 && isfinite(p[0]) && p[1] > p[0] && isfinite(p[1]) && p[2] >= 0.0 &&
               p[2] <= 1.0;
    return 0;
}

/* psi(x), x > 0: the recurrence psi(x) = psi(x + 1) - 1/x up to x >= 6, then the asymptotic
 * series ln x - 1/(2x) - sum B_2k / (2k x^2k) (Abramowitz & Stegun 6.3.18). */
static double digamma(double x) {
    double acc = 0.0;
    while (x < 6.0) {
        acc -= 1.0 / x;
        x += 1.0;
    }
    double q = 1.0 / (x * x);
    double series = q * (1.0 / 12.0 - q * (1.0 / 120.0 - q * (1.0 / 252.0 - q * (1.0 / 240.0 -
                                                                                 q * (1.0 / 132.0)))));
    return acc + log(x) - 0.5 / x - series;
}

/* The modified gamma law in y = b r^gamma: the number density is Gamma((alpha + 1)/gamma) in
 * y, the area-weighted one Gamma(kA = (alpha + 3)/gamma). The range bounds both tails of the
 * area-weighted law (the cross sections'): below, P(Y < y) <= y^kA / Gamma(kA + 1) = e^-G,
 * ln y_lo = (ln Gamma(kA + 1) - G)/kA; above, as for the gamma law (Chernoff),
 * y_hi = kA + G + sqrt(2 G kA); t = (ln y - ln b)/gamma. Number averages are those of the law
 * truncated to the range. With the derivatives of (t_lo, t_hi) with respect to the parameters
 * when dlo, dhi are not NULL. */
static void mgamma_range(const double *p, double *t_lo, double *t_hi, double dlo[3],
                         double dhi[3]) {
    const double G = GAMMA_TAIL_MG;
    double alpha = p[0], b = p[1], g = p[2];
    double ka = (alpha + 3.0) / g;
    double lylo = (lgamma(ka + 1.0) - G) / ka;
    double rt = sqrt(2.0 * G * ka);
    double yhi = ka + G + rt;
    double lyhi = log(yhi);
    *t_lo = (lylo - log(b)) / g;
    *t_hi = (lyhi - log(b)) / g;
    if (dlo) {
        double dlylo_dka = (digamma(ka + 1.0) - lylo) / ka;
        double dlyhi_dka = (1.0 + G / rt) / yhi;
        /* dkA/dalpha = 1/gamma, dkA/dgamma = -kA/gamma */
        dlo[0] = dlylo_dka / (g * g);
        dlo[1] = -1.0 / (b * g);
        dlo[2] = dlylo_dka * (-ka / g) / g - *t_lo / g;
        dhi[0] = dlyhi_dka / (g * g);
        dhi[1] = -1.0 / (b * g);
        dhi[2] = dlyhi_dka * (-ka / g) / g - *t_hi / g;
    }
}

/* The fixed radius range in ln r. Lognormal: ln r_g -+ 5 ln sigma_g. Gamma (a = r_eff,
 * b = v_eff): from ln a - 8 s, s = sqrt(b / (1 - 2b)) the reference's log-width of the gamma
 * peak (distribution.rs:319-323), the lower side long for the left skew of ln r under a gamma
 * law; to ln a + ln(1 + G b + sqrt(2 G b)), G = 37: the area-weighted law is Gamma(k = 1/b,
 * scale a b) in r, whose upper tail beyond k + c scale units is at most
 * exp(-(c - k ln(1 + c/k))) (Chernoff), below e^-G for c = G + sqrt(2 k G). Modified gamma:
 * mgamma_range. Power law: its support [ln r_min, ln r_max]. Monodisperse: the radius (a range
 * of width 0). */
static const double GAMMA_TAIL = 37.0;
int orc_radius_range(int kind, const double *p, double *t_lo, double *t_hi) {
    if (!dist_valid(kind, p)) return -1;
    if (kind == ORC_LOGNORMAL) {
        double c = log(p[0]), h = 5.0 * log(p[1]);
        *t_lo = c - h;
        *t_hi = c + h;
    } else if (kind == ORC_GAMMA) {
        double s = sqrt(p[1] / (1.0 - 2.0 * p[1]));
        double c = log(p[0]);
        *t_lo = c - 8.0 * s;
        *t_hi = c + log(1.0 + GAMMA_TAIL * p[1] + sqrt(2.0 * GAMMA_TAIL * p[1]));
    } else if (kind == ORC_MODIFIED_GAMMA) {
        mgamma_range(p, t_lo, t_hi, NULL, NULL);
    } else if (kind == ORC_POWER_LAW) {
        *t_lo = log(p[1]);
        *t_hi = log(p[2]);
    } else if (kind == ORC_SEGMENT) {
        *t_lo = log(p[0]);
        *t_hi = log(p[1]);
    } else {
        *t_lo = *t_hi = log(p[0]);
    }
    return 0;
}

/* The quadrature weight of radius node i: w_i (t_hi - t_lo)/2 u(r) r, and w_i r for a
 * monodisperse law (all nodes at its radius). */
static double node_weight(int kind, double w, double half, double u, double r) {
    return w * (kind == ORC_MONODISPERSE ? 1.0 : half) * u * r;
}

/* ---- the fixed-node ensemble ---- */

/* The outputs of an ensemble from its sums S (8) and summed tables T (4 x n_angles, divided by
 * Z on return) (ensemble/engine.rs:1782-1836,2396-2497). */
static void ensemble_outputs(const double S[8], double *T, int64_t n_angles, const double *mu,
                             const double *mw, int64_t order, int normalized, double *scalars,
                             double *greek, double *phase) {
    double z = S[0];
    double ext = S[1] / z, sca = S[2] / z, scag = S[3] / z, absb = S[4] / z, back = S[5] / z;
    double r2 = S[6] / z, r3 = S[7] / z;
    scalars[0] = ext;
    scalars[1] = sca;
    scalars[2] = absb;
    scalars[3] = back;
    scalars[4] = sca / ext;
    scalars[5] = scag / sca;
    scalars[6] = 4.0 * PI * (ext / back);
    scalars[7] = 0.0;
    scalars[8] = r3 / r2;
    scalars[9] = PI * r2;
    scalars[10] = 4.0 * PI / 3.0 * r3;
    for (int64_t j = 0; j < 4 * n_angles; j++) T[j] /= z;
    if (greek && order >= 0) {
        orc_greek_project(n_angles, mu, mw, T, T + n_angles, T + 2 * n_angles, T + 3 * n_angles,
                          order, greek);
        if (normalized) {
            double a0 = greek[0];
            for (int64_t l = 0; l < 6 * (order + 1); l++) greek[l] /= a0;
            greek[0] = 1.0;
        }
    }
    if (phase) {
        int src[6] = {0, 1, 0, 2, 3, 2};
        for (int e = 0; e < 6; e++)
            for (int64_t j = 0; j < n_angles; j++) {
                double v = T[src[e] * n_angles + j];
                phase[e * n_angles + j] = normalized ? v / sca : v;
            }
    }
}

int orc_ensemble_fixed_ex(int kind, const double *p, double lambda, int layers,
                          const double *fractions, const orc_cx *m, int64_t n_radius,
                          int64_t n_angles, int64_t order, int normalized, const double *range,
                          const int64_t *n_nodes, double *scalars, double *greek,
                          double *phase) {
    return orc_ensemble_fixed_mu(kind, p, lambda, layers, fractions, m, n_radius, n_angles, order,
                                 normalized, range, n_nodes, NULL, 0, scalars, greek, phase);
}

int orc_ensemble_fixed_mu(int kind, const double *p, double lambda, int layers,
                          const double *fractions, const orc_cx *m, int64_t n_radius,
                          int64_t n_angles, int64_t order, int normalized, const double *range,
                          const int64_t *n_nodes, const double *mu_user, int64_t n_mu,
                          double *scalars, double *greek, double *phase) {
    double t_lo, t_hi;
    if (orc_radius_range(kind, p, &t_lo, &t_hi) != 0) return -1;
    if (range) {
        t_lo = range[0];
        t_hi = range[1];
    }
    if (!(lambda > 0.0 && isfinite(lambda)) || layers < 1 || n_radius < 1 || n_angles < 1) return -2;
    double *rn = (double *)malloc(sizeof(double) * (size_t)n_radius);
    double *rw = (double *)malloc(sizeof(double) * (size_t)n_radius);
    orc_gauss_legendre(n_radius, rn, rw);
    double *mu = (double *)malloc(sizeof(double) * (size_t)n_angles);
    double *mw = (double *)malloc(sizeof(double) * (size_t)n_angles);
    orc_gauss_legendre(n_angles, mu, mw);
    double half = 0.5 * (t_hi - t_lo), mid = 0.5 * (t_hi + t_lo);
    double S[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double *T = (double *)calloc(4 * (size_t)n_angles, sizeof(double));
    double *xl = (double *)malloc(sizeof(double) * (size_t)layers);
    cx *s1 = (cx *)malloc(sizeof(cx) * (size_t)n_angles);
    cx *s2 = (cx *)malloc(sizeof(cx) * (size_t)n_angles);
    /* the phase tables at the caller's cosines, when given */
    const int64_t nu = mu_user ? n_mu : 0;
    double *Tu = (double *)calloc(4 * (size_t)(nu > 0 ? nu : 1), sizeof(double));
    cx *u1 = (cx *)malloc(sizeof(cx) * (size_t)(nu > 0 ? nu : 1));
    cx *u2 = (cx *)malloc(sizeof(cx) * (size_t)(nu > 0 ? nu : 1));
    for (int64_t i = 0; i < n_radius; i++) {
        double t = mid + half * rn[i];
        double r = exp(t);
        double u = orc_density(kind, p, r, NULL);
        double W = node_weight(kind, rw[i], half, u, r);
        double x = 2.0 * PI * r / lambda;
        int64_t n = n_nodes ? n_nodes[i] : orc_default_n_max(x);
        cx *ab = (cx *)malloc(sizeof(cx) * 2 * (size_t)n);
        double abs_sum;
        if (layers == 1) {
            double aj[3];
            orc_mie_solve(x, m[0], n, 0, ab, ab + n, NULL, NULL, NULL, NULL, &abs_sum, aj);
        } else {
            for (int l = 0; l < layers - 1; l++) xl[l] = fractions[l] * x;
            xl[layers - 1] = x;
            orc_layered_solve(layers, xl, m, n, 0, ab, ab + n, NULL, &abs_sum, NULL);
        }
        double q[5];
        orc_efficiencies(x, n, ab, ab + n, abs_sum, q);
        double area = PI * r * r;
        double g[8] = {1.0, area * q[0], area * q[1], area * q[1] * q[4], area * q[2],
                       area * q[3], r * r, r * r * r};
        for (int k = 0; k < 8; k++) S[k] += W * g[k];
        orc_amplitudes(n, ab, ab + n, mu, n_angles, s1, s2);
        double c = area * (4.0 / (x * x));
        for (int64_t j = 0; j < n_angles; j++) {
            double n1 = norm2(s1[j]), n2 = norm2(s2[j]);
            cx pp = mul(s2[j], C(s1[j].re, -s1[j].im));
            T[j] += W * (c * (0.5 * (n1 + n2)));
            T[n_angles + j] += W * (c * (0.5 * (n2 - n1)));
            T[2 * n_angles + j] += W * (c * pp.re);
            T[3 * n_angles + j] += W * (c * pp.im);
        }
        if (nu > 0) {
            orc_amplitudes(n, ab, ab + n, mu_user, nu, u1, u2);
            for (int64_t j = 0; j < nu; j++) {
                double n1 = norm2(u1[j]), n2 = norm2(u2[j]);
                cx pp = mul(u2[j], C(u1[j].re, -u1[j].im));
                Tu[j] += W * (c * (0.5 * (n1 + n2)));
                Tu[nu + j] += W * (c * (0.5 * (n2 - n1)));
                Tu[2 * nu + j] += W * (c * pp.re);
                Tu[3 * nu + j] += W * (c * pp.im);
            }
        }
        free(ab);
    }
    ensemble_outputs(S, T, n_angles, mu, mw, order, normalized, scalars, greek,
                     nu > 0 ? NULL : phase);
    if (nu > 0 && phase) {
        /* [F11, F12, F22 = F11, F33, F34, F44 = F33] at the caller's cosines, divided by Z and,
         * normalized, by <C_sca> (engine.rs:2459-2490) */
        double z = S[0], sca = S[2] / z;
        int src[6] = {0, 1, 0, 2, 3, 2};
        for (int e = 0; e < 6; e++)
            for (int64_t j = 0; j < nu; j++) {
                double v = Tu[src[e] * nu + j] / z;
                phase[e * nu + j] = normalized ? v / sca : v;
            }
    }
    free(rn); free(rw); free(mu); free(mw); free(T); free(xl); free(s1); free(s2);
    free(Tu); free(u1); free(u2);
    return 0;
}

int orc_ensemble_fixed(int kind, const double *p, double lambda, int layers,
                       const double *fractions, const orc_cx *m, int64_t n_radius,
                       int64_t n_angles, int64_t order, int normalized, double *scalars,
                       double *greek, double *phase) {
    return orc_ensemble_fixed_ex(kind, p, lambda, layers, fractions, m, n_radius, n_angles,
                                 order, normalized, NULL, NULL, scalars, greek, phase);
}

/* d(ln u)/d(ln r) at fixed parameters, from the densities of distribution.rs:366-430. */
static double log_density_slope(int kind, const double *p, double r) {
    if (kind == ORC_LOGNORMAL) {
        double s = log(p[1]);
        return -(log(r) - log(p[0])) / (s * s) - 1.0;
    }
    if (kind == ORC_GAMMA) return (1.0 - 3.0 * p[1]) / p[1] - r / (p[0] * p[1]);
    if (kind == ORC_MODIFIED_GAMMA) return p[0] - p[1] * p[2] * pow(r, p[2]);
    if (kind == ORC_POWER_LAW) return -p[0];
    if (kind == ORC_SEGMENT) {
        double w = p[1] - p[0], s = (r - p[0]) / w;
        return r * (2.0 * p[2] - 1.0) / (w * ((1.0 - p[2]) * (1.0 - s) + p[2] * s));
    }
    return 0.0;
}

/* Derivatives of the range's midpoint and half-width in ln r with respect to the three
 * parameters; returns the half-width. */
static double range_derivatives(int kind, const double *p, double dmid[3], double dhalf[3]) {
    for (int k = 0; k < 3; k++) dmid[k] = dhalf[k] = 0.0;
    if (kind == ORC_LOGNORMAL) {
        dmid[0] = 1.0 / p[0];
        dhalf[1] = 5.0 / p[1];
        return 5.0 * log(p[1]);
    }
    if (kind == ORC_GAMMA) {
        double q = 1.0 - 2.0 * p[1], s = sqrt(p[1] / q);
        double ds = 0.5 / (s * q * q);
        double rt = sqrt(2.0 * GAMMA_TAIL * p[1]);
        double up = log(1.0 + GAMMA_TAIL * p[1] + rt);
        double dup = (GAMMA_TAIL + GAMMA_TAIL / rt) / (1.0 + GAMMA_TAIL * p[1] + rt);
        dmid[0] = 1.0 / p[0];
        dmid[1] = 0.5 * (dup - 8.0 * ds);
        dhalf[1] = 0.5 * (dup + 8.0 * ds);
        return 0.5 * (up + 8.0 * s);
    }
    if (kind == ORC_MODIFIED_GAMMA) {
        double t_lo, t_hi, dlo[3], dhi[3];
        mgamma_range(p, &t_lo, &t_hi, dlo, dhi);
        for (int k = 0; k < 3; k++) {
            dmid[k] = 0.5 * (dhi[k] + dlo[k]);
            dhalf[k] = 0.5 * (dhi[k] - dlo[k]);
        }
        return 0.5 * (t_hi - t_lo);
    }
    if (kind == ORC_POWER_LAW) return 0.5 * (log(p[2]) - log(p[1]));
    if (kind == ORC_SEGMENT) return 0.5 * (log(p[1]) - log(p[0]));
    dmid[0] = 1.0 / p[0]; /* monodisperse: every node at the radius */
    return 0.0;
}

/* The derivatives of the phase elements from those of S1, S2 (mie/phase.rs:60-73). */
static void element_derivatives(cx s1, cx s2, cx d1, cx d2, double out[4]) {
    double r1 = re_mul_conj(s1, d1), r2 = re_mul_conj(s2, d2);
    cx p = add(mul(d2, C(s1.re, -s1.im)), mul(s2, C(d1.re, -d1.im)));
    out[0] = r1 + r2;
    out[1] = r2 - r1;
    out[2] = p.re;
    out[3] = p.im;
}

/* Forward-mode derivatives of the fixed path, direction by direction, as the reference forms
 * its ensemble derivatives: per node the efficiency Jacobians of every layer
 * (efficiencies.rs:192-230, layered/kernel.rs:549-587), the radius derivative at fixed
 * fractions 2 pi r Q + pi r^2 (2 pi / lambda) dQ/dx (engine.rs:1111-1123, particle.rs:221-249),
 * the phase-element derivatives from the amplitude derivatives (phase.rs:60-73); the weights'
 * derivatives; then the quotient rule through the averages, the Greek normalization and the
 * phase normalization (engine.rs:1782-1836,2396-2497; lin.rs:7-68). Directions
 * P = 3 + 2L + (L - 1): [p0, p1, p2, Re m_0, Im m_0, ..., fraction_0 ...] (the columns of a
 * distribution's unused or fixed parameters are 0). With fixed_range the ln r range is held
 * fixed (the parameters act through the density alone, the reference's definition); otherwise
 * the nodes move with the parameters. n_nodes (may be NULL) receives N per node. Returns 0 or a
 * negative code. */
int orc_ensemble_fixed_jacobian(int kind, const double *p, double lambda, int layers,
                                const double *fractions, const orc_cx *m, int64_t n_radius,
                                int64_t n_angles, int64_t order, int normalized,
                                int fixed_range, double *scalars, double *greek, double *phase,
                                double *d_scalars, double *d_greek, double *d_phase,
                                int64_t *n_nodes) {
    double t_lo, t_hi;
    if (orc_radius_range(kind, p, &t_lo, &t_hi) != 0) return -1;
    if (!(lambda > 0.0 && isfinite(lambda)) || layers < 1 || n_radius < 1 || n_angles < 1) return -2;
    const int L = layers;
    const int P = 3 + 2 * L + (L - 1);
    double *rn = (double *)malloc(sizeof(double) * (size_t)n_radius);
    double *rw = (double *)malloc(sizeof(double) * (size_t)n_radius);
    orc_gauss_legendre(n_radius, rn, rw);
    double *mu = (double *)malloc(sizeof(double) * (size_t)n_angles);
    double *mw = (double *)malloc(sizeof(double) * (size_t)n_angles);
    orc_gauss_legendre(n_angles, mu, mw);
    double half = 0.5 * (t_hi - t_lo), mid = 0.5 * (t_hi + t_lo);
    double dmid[3], dhalf[3];
    range_derivatives(kind, p, dmid, dhalf);
    double S[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    double *dS = (double *)calloc(8 * (size_t)P, sizeof(double));
    double *T = (double *)calloc(4 * (size_t)n_angles, sizeof(double));
    double *dT = (double *)calloc(4 * (size_t)n_angles * (size_t)P, sizeof(double));
    double *xl = (double *)malloc(sizeof(double) * (size_t)L);
    cx *s1 = (cx *)malloc(sizeof(cx) * (size_t)n_angles);
    cx *s2 = (cx *)malloc(sizeof(cx) * (size_t)n_angles);
    cx *ds1 = (cx *)malloc(sizeof(cx) * (size_t)n_angles);
    cx *ds2 = (cx *)malloc(sizeof(cx) * (size_t)n_angles);
    double *jac = (double *)malloc(sizeof(double) * 15 * (size_t)L);
    double *abs_jac = (double *)malloc(sizeof(double) * 3 * (size_t)L);
    for (int64_t i = 0; i < n_radius; i++) {
        double t = mid + half * rn[i];
        double r = exp(t);
        double ugrad[3];
        double u = orc_density(kind, p, r, ugrad);
        double W = node_weight(kind, rw[i], half, u, r);
        double x = 2.0 * PI * r / lambda;
        int64_t n = orc_default_n_max(x);
        if (n_nodes) n_nodes[i] = n;
        cx *ab = (cx *)malloc(sizeof(cx) * 2 * (size_t)n);
        cx *a = ab, *b = ab + n;
        cx *dg = (cx *)malloc(sizeof(cx) * 4 * (size_t)n * (size_t)L); /* [l][da_dx, da_dm, db_dx, db_dm][k] */
        cx *da = (cx *)malloc(sizeof(cx) * 2 * (size_t)n), *db = da + n;
        double abs_sum;
        if (L == 1) {
            orc_mie_solve(x, m[0], n, 1, a, b, dg, dg + n, dg + 2 * n, dg + 3 * n, &abs_sum, abs_jac);
        } else {
            for (int l = 0; l < L - 1; l++) xl[l] = fractions[l] * x;
            xl[L - 1] = x;
            orc_layered_solve(L, xl, m, n, 1, a, b, dg, &abs_sum, abs_jac);
        }
        double q[5];
        orc_efficiencies(x, n, a, b, abs_sum, q);
        for (int l = 0; l < L; l++) {
            const cx *g = dg + (size_t)l * 4 * (size_t)n;
            orc_efficiencies_jacobian(x, n, a, b, g, g + n, g + 2 * n, g + 3 * n, abs_sum,
                                      abs_jac + 3 * l, l == L - 1, jac + 15 * l);
        }
        double area = PI * r * r;
        double h[8] = {1.0, area * q[0], area * q[1], area * q[1] * q[4], area * q[2],
                       area * q[3], r * r, r * r * r};
        for (int k = 0; k < 8; k++) S[k] += W * h[k];
        orc_amplitudes(n, a, b, mu, n_angles, s1, s2);
        double c = area * (4.0 / (x * x));
        for (int64_t j = 0; j < n_angles; j++) {
            double n1 = norm2(s1[j]), n2 = norm2(s2[j]);
            cx pp = mul(s2[j], C(s1[j].re, -s1[j].im));
            T[j] += W * (c * (0.5 * (n1 + n2)));
            T[n_angles + j] += W * (c * (0.5 * (n2 - n1)));
            T[2 * n_angles + j] += W * (c * pp.re);
            T[3 * n_angles + j] += W * (c * pp.im);
        }
        /* Directions of the particle: d = 0 the size at fixed fractions (x d/dx), then
         * Re m_l, Im m_l per layer, then the fractions (x d/dx_l). */
        const int ND = 1 + 2 * L + (L - 1);
        for (int d = 0; d < ND; d++) {
            double dq[5] = {0, 0, 0, 0, 0};
            for (int64_t k = 0; k < n; k++) {
                da[k] = R(0.0);
                db[k] = R(0.0);
            }
            if (d == 0) {
                for (int l = 0; l < L; l++) {
                    double f = (l == L - 1) ? x : fractions[l] * x;
                    const cx *g = dg + (size_t)l * 4 * (size_t)n;
                    for (int e = 0; e < 5; e++) dq[e] += f * jac[15 * l + e];
                    for (int64_t k = 0; k < n; k++) {
                        da[k] = add(da[k], scl(g[k], f));
                        db[k] = add(db[k], scl(g[2 * n + k], f));
                    }
                }
            } else if (d <= 2 * L) {
                int l = (d - 1) / 2, im = (d - 1) % 2;
                const cx *g = dg + (size_t)l * 4 * (size_t)n;
                for (int e = 0; e < 5; e++) dq[e] = jac[15 * l + 5 * (1 + im) + e];
                for (int64_t k = 0; k < n; k++) {
                    da[k] = im ? times_i(g[n + k]) : g[n + k];
                    db[k] = im ? times_i(g[3 * n + k]) : g[3 * n + k];
                }
            } else {
                int l = d - 1 - 2 * L;
                const cx *g = dg + (size_t)l * 4 * (size_t)n;
                for (int e = 0; e < 5; e++) dq[e] = x * jac[15 * l + e];
                for (int64_t k = 0; k < n; k++) {
                    da[k] = scl(g[k], x);
                    db[k] = scl(g[2 * n + k], x);
                }
            }
            /* the node's per-particle quantities along this direction */
            double dh[8] = {0.0, area * dq[0], area * dq[1], area * (dq[1] * q[4] + q[1] * dq[4]),
                            area * dq[2], area * dq[3], 0.0, 0.0};
            orc_amplitudes(n, da, db, mu, n_angles, ds1, ds2);
            if (d == 0) {
                /* d/d(ln r) at fixed fractions: the explicit r^2 of pi r^2 and r^2, r^3 */
                for (int k = 1; k <= 5; k++) dh[k] += 2.0 * h[k];
                dh[6] = 2.0 * h[6];
                dh[7] = 3.0 * h[7];
                /* A monodisperse law's range is its radius: the radius moves its nodes
                 * under either definition. */
                const int hold = fixed_range && kind != ORC_MONODISPERSE;
                for (int th = 0; th < 3; th++) {
                    double dlnr = hold ? 0.0 : dmid[th] + rn[i] * dhalf[th];
                    double dlnw = (hold || half == 0.0 ? 0.0 : dhalf[th] / half) +
                                  ugrad[th] / u + (log_density_slope(kind, p, r) + 1.0) * dlnr;
                    double *dSp = dS + 8 * th;
                    for (int k = 0; k < 8; k++) dSp[k] += W * dlnw * h[k] + W * dh[k] * dlnr;
                    double *dTp = dT + (size_t)th * 4 * (size_t)n_angles;
                    for (int64_t j = 0; j < n_angles; j++) {
                        double f[4], df[4];
                        double n1 = norm2(s1[j]), n2 = norm2(s2[j]);
                        cx pp = mul(s2[j], C(s1[j].re, -s1[j].im));
                        f[0] = 0.5 * (n1 + n2);
                        f[1] = 0.5 * (n2 - n1);
                        f[2] = pp.re;
                        f[3] = pp.im;
                        element_derivatives(s1[j], s2[j], ds1[j], ds2[j], df);
                        for (int e = 0; e < 4; e++)
                            dTp[e * n_angles + j] += W * c * (dlnw * f[e] + df[e] * dlnr);
                    }
                }
            } else {
                int pc = d + 2;
                double *dSp = dS + 8 * pc;
                for (int k = 0; k < 8; k++) dSp[k] += W * dh[k];
                double *dTp = dT + (size_t)pc * 4 * (size_t)n_angles;
                for (int64_t j = 0; j < n_angles; j++) {
                    double df[4];
                    element_derivatives(s1[j], s2[j], ds1[j], ds2[j], df);
                    for (int e = 0; e < 4; e++) dTp[e * n_angles + j] += W * c * df[e];
                }
            }
        }
        free(ab);
        free(dg);
        free(da);
    }
    /* values */
    double *Tv = (double *)malloc(sizeof(double) * 4 * (size_t)n_angles);
    memcpy(Tv, T, sizeof(double) * 4 * (size_t)n_angles);
    ensemble_outputs(S, Tv, n_angles, mu, mw, order, normalized, scalars, greek, phase);
    /* the quotient rule per direction (Tv now holds F = T / Z) */
    double z = S[0];
    double qv[8];
    for (int k = 0; k < 8; k++) qv[k] = S[k] / z;
    const double ext = qv[1], sca = qv[2], scag = qv[3], back = qv[5], r2 = qv[6], r3 = qv[7];
    const int64_t len = order + 1;
    double *dF = (double *)malloc(sizeof(double) * 4 * (size_t)n_angles);
    double *dgam = (greek && d_greek && order >= 0) ? (double *)malloc(sizeof(double) * 6 * (size_t)len) : NULL;
    double *gam = dgam ? (double *)malloc(sizeof(double) * 6 * (size_t)len) : NULL;
    if (gam) orc_greek_project(n_angles, mu, mw, Tv, Tv + n_angles, Tv + 2 * n_angles,
                               Tv + 3 * n_angles, order, gam);
    for (int p = 0; p < P; p++) {
        const double *dSp = dS + 8 * p;
        double dz = dSp[0];
        double dq[8];
        for (int k = 1; k < 8; k++) dq[k] = (dSp[k] - qv[k] * dz) / z;
        double *ds = d_scalars + 11 * p;
        ds[0] = dq[1];
        ds[1] = dq[2];
        ds[2] = dq[4];
        ds[3] = dq[5];
        ds[4] = (dq[2] - (sca / ext) * dq[1]) / ext;
        ds[5] = (dq[3] - (scag / sca) * dq[2]) / sca;
        ds[6] = 4.0 * PI * (dq[1] * back - ext * dq[5]) / (back * back);
        ds[7] = 0.0;
        ds[8] = (dq[7] - (r3 / r2) * dq[6]) / r2;
        ds[9] = PI * dq[6];
        ds[10] = 4.0 * PI / 3.0 * dq[7];
        const double *dTp = dT + (size_t)p * 4 * (size_t)n_angles;
        for (int64_t j = 0; j < 4 * n_angles; j++) dF[j] = (dTp[j] - Tv[j] * dz) / z;
        if (dgam) {
            orc_greek_project(n_angles, mu, mw, dF, dF + n_angles, dF + 2 * n_angles,
                              dF + 3 * n_angles, order, dgam);
            double *dg = d_greek + (size_t)p * 6 * (size_t)len;
            if (normalized) {
                double a0 = gam[0], da0 = dgam[0];
                for (int64_t l = 0; l < 6 * len; l++) dg[l] = (dgam[l] - (gam[l] / a0) * da0) / a0;
                dg[0] = 0.0;
            } else {
                for (int64_t l = 0; l < 6 * len; l++) dg[l] = dgam[l];
            }
        }
        if (d_phase) {
            int src[6] = {0, 1, 0, 2, 3, 2};
            double *dp = d_phase + (size_t)p * 6 * (size_t)n_angles;
            for (int e = 0; e < 6; e++)
                for (int64_t j = 0; j < n_angles; j++) {
                    double f = Tv[src[e] * n_angles + j], df = dF[src[e] * n_angles + j];
                    dp[e * n_angles + j] = normalized ? (df - (f / sca) * dq[2]) / sca : df;
                }
        }
    }
    free(rn); free(rw); free(mu); free(mw); free(dS); free(T); free(dT); free(xl);
    free(s1); free(s2); free(ds1); free(ds2); free(jac); free(abs_jac); free(Tv); free(dF);
    free(dgam); free(gam);
    return 0;
}

int orc_ensemble_batch(int64_t count, const int *kind, const double *p0, const double *p1,
                       const double *p2, const double *lambda, const double *m, int64_t n_radius,
                       int64_t n_angles, int64_t order, int normalized, double *scalars,
                       double *greek, double *phase, int threads) {
    int bad = 0;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) reduction(| : bad)
#endif
    for (int64_t e = 0; e < count; e++) {
        cx mm = C(m[2 * e], m[2 * e + 1]);
        const double pe[3] = {p0[e], p1[e], p2 ? p2[e] : 0.0};
        int rc = orc_ensemble_fixed(kind[e], pe, lambda[e], 1, NULL, &mm, n_radius,
                                    n_angles, order, normalized, scalars + 11 * e,
                                    greek ? greek + 6 * (order + 1) * e : NULL,
                                    phase ? phase + 6 * n_angles * e : NULL);
        if (rc != 0) bad = 1;
    }
    (void)threads;
    return bad ? -1 : 0;
}
