// One sphere per 64-lane wavefront, for batches too small to fill the GPU with one sphere per
// lane (where a launch lasts as long as its largest sphere's serial recurrences).
//
// Every recurrence of the solve is split into 64 contiguous chunks, one per lane:
//  * the downward ratio chains u_k = 1 / (c_k - u_{k+1}) of m x and of x (c_k = (2k + 1) / z)
//    are Moebius maps, [[0, 1], [-1, c_k]] acting on homogeneous pairs (p, q), u = p / q;
//  * psi_k, chi_k below floor(x) follow f_k = c f_{k-1} - f_{k-2} (c = (2k - 1) / x), the
//    linear map [[c, -1], [1, 0]] on (f_{k-1}, f_{k-2});
//  * above floor(x) the ratio r_k = 1 / ((2k - 1) / x - r_{k-1}) is again a Moebius map, and
//    t_k = t_{k-1} r_k u_k(x), 1/xi_k = 1/xi_{k-1} r_k are running products.
// Each lane composes its chunk's 2x2 matrix (scaled by powers of two, exact, for the projective
// maps), a six-step shuffle scan gives every lane the composition of the chunks before its own,
// and the lane applies it to the recurrence's start and replays its chunk with the serial
// solver's own checked steps (step_complex, step_real, inv), writing the chains' values to the
// wavefront's columns. The order loop (mie_orders) then runs on each lane's range of orders
// from its start state, and the efficiency sums are added over the lanes in a fixed tree.
// The composition is exact algebra; its rounding enters only at the chunk boundaries, where it
// is that of a product of a few dozen 2x2 matrices.
#pragma once

#include "mie_device.h"

namespace scatrs {

template <class T>
struct Mat2 {
    T a, b, c, d;  // [[a, b], [c, d]]
};

SCATRS_HD double scalar_one(double) { return 1.0; }
SCATRS_HD cx scalar_one(cx) { return real(1.0); }
SCATRS_HD double scalar_zero(double) { return 0.0; }
SCATRS_HD cx scalar_zero(cx) { return real(0.0); }

template <class T>
__device__ __forceinline__ Mat2<T> mat_identity() {
    const T o = scalar_one(T{}), z = scalar_zero(T{});
    return Mat2<T>{o, z, z, o};
}

// X Y (Y applied first).
template <class T>
__device__ __forceinline__ Mat2<T> mat_mul(const Mat2<T> &X, const Mat2<T> &Y) {
    return Mat2<T>{X.a * Y.a + X.b * Y.c, X.a * Y.b + X.b * Y.d, X.c * Y.a + X.d * Y.c,
                   X.c * Y.b + X.d * Y.d};
}

__device__ __forceinline__ int exp_of(double v) {
    return v == 0.0 ? -4000 : __builtin_amdgcn_frexp_exp(v);
}
__device__ __forceinline__ int exp_of(cx v) { return max(exp_of(v.re), exp_of(v.im)); }
__device__ __forceinline__ double scale2(double v, int e) { return __builtin_amdgcn_ldexp(v, e); }
__device__ __forceinline__ cx scale2(cx v, int e) {
    return cx{__builtin_amdgcn_ldexp(v.re, e), __builtin_amdgcn_ldexp(v.im, e)};
}

// The projective map M scaled by a power of two so that its largest entry lies in [0.5, 1).
template <class T>
__device__ __forceinline__ Mat2<T> mat_normalized(const Mat2<T> &M) {
    const int e = max(max(exp_of(M.a), exp_of(M.b)), max(exp_of(M.c), exp_of(M.d)));
    if (e == -4000) return M;
    return Mat2<T>{scale2(M.a, -e), scale2(M.b, -e), scale2(M.c, -e), scale2(M.d, -e)};
}

__device__ __forceinline__ double shfl_up_v(double v, int d) { return __shfl_up(v, d, 64); }
__device__ __forceinline__ cx shfl_up_v(cx v, int d) {
    return cx{__shfl_up(v.re, d, 64), __shfl_up(v.im, d, 64)};
}
__device__ __forceinline__ double shfl_v(double v, int l) { return __shfl(v, l, 64); }
__device__ __forceinline__ cx shfl_v(cx v, int l) {
    return cx{__shfl(v.re, l, 64), __shfl(v.im, l, 64)};
}

template <class T>
__device__ __forceinline__ Mat2<T> mat_shfl_up(const Mat2<T> &M, int d) {
    return Mat2<T>{shfl_up_v(M.a, d), shfl_up_v(M.b, d), shfl_up_v(M.c, d), shfl_up_v(M.d, d)};
}
template <class T>
__device__ __forceinline__ Mat2<T> mat_shfl(const Mat2<T> &M, int l) {
    return Mat2<T>{shfl_v(M.a, l), shfl_v(M.b, l), shfl_v(M.c, l), shfl_v(M.d, l)};
}

// The scan over the lanes of chunk maps M (lane 0's applied first): returns the composition of
// the chunks before this lane's (the identity on lane 0) and sets *all to the composition of
// every chunk. PROJ: scaled after every product (projective maps).
template <bool PROJ, class T>
__device__ __forceinline__ Mat2<T> scan_maps(Mat2<T> M, int lane, Mat2<T> *all) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const Mat2<T> o = mat_shfl_up(M, d);
        if (lane >= d) {
            M = mat_mul(M, o);
            if constexpr (PROJ) M = mat_normalized(M);
        }
    }
    if (all) *all = mat_shfl(M, 63);
    Mat2<T> e = mat_shfl_up(M, 1);
    if (lane == 0) e = mat_identity<T>();
    return e;
}

// The same for running products of complex factors.
__device__ __forceinline__ cx scan_products(cx v, int lane) {
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const cx o = shfl_up_v(v, d);
        if (lane >= d) v = v * o;
    }
    cx e = shfl_up_v(v, 1);
    if (lane == 0) e = real(1.0);
    return e;
}

// Lane `lane`'s chunk of count items split into 64 contiguous chunks: items [*lo, *hi) of
// 0..count (empty when *lo >= *hi).
__device__ __forceinline__ void lane_chunk(int count, int lane, int *lo, int *hi) {
    const int c = (count + 63) >> 6;
    *lo = min(count, lane * c);
    *hi = min(count, *lo + c);
}

// The downward ratio chain u_k = 1 / (c_k - u_{k+1}) for k = top .. bottom from u_{top+1} = 0,
// its values for k <= keep given to store(k, u): coef(k) gives c_k and step(k, u) the serial
// solver's step from u_{k+1} to u_k (T = double or cx).
template <class T, class Coef, class Step, class Store>
__device__ __forceinline__ void chain_down(int top, int bottom, int keep, Coef coef, Step step,
                                           Store store, int lane) {
    if (top < bottom) return;
    int lo, hi;
    lane_chunk(top - bottom + 1, lane, &lo, &hi);
    // Lane chunk: indices top - lo down to top - hi + 1.
    Mat2<T> M = mat_identity<T>();
    for (int i = lo; i < hi; ++i) {
        const T c = coef(top - i);
        M = Mat2<T>{M.c, M.d, c * M.c - M.a, c * M.d - M.b};
        if (((i - lo) & 3) == 3) M = mat_normalized(M);
    }
    M = mat_normalized(M);
    const Mat2<T> E = scan_maps<true>(M, lane, static_cast<Mat2<T> *>(nullptr));
    // (p, q) = E (0, 1): u_{top - lo + 1} = b / d.
    T u;
    if constexpr (sizeof(T) == sizeof(double)) {
        u = lo == 0 ? 0.0 : E.b / E.d;
    } else {
        u = lo == 0 ? real(0.0) : E.b * inv(E.d);
    }
    for (int i = lo; i < hi; ++i) {
        const int k = top - i;
        u = step(k, u);
        if (k <= keep) store(k, u);
    }
}

// What a lane does with each of its orders' coefficients besides the efficiency sums: nothing,
// or write them to a node's staging rows (FP64 a_1, b_1, a_2, ...; or one fp32 float4 (a.re,
// a.im, b.re, b.im) per order), as the ensembles' node solves stage them.
struct NoStage {
    __device__ void operator()(int, cx, cx) const {}
};
struct StageF64 {
    cx *ab;
    __device__ void operator()(int k, cx a, cx b) const {
        ab[2 * (k - 1)] = a;
        ab[2 * (k - 1) + 1] = b;
    }
};
struct StageF32 {
    float4 *ab;
    __device__ void operator()(int k, cx a, cx b) const {
        ab[k - 1] = make_float4((float)a.re, (float)a.im, (float)b.re, (float)b.im);
    }
};

// The efficiency sink of one lane's orders: also keeps the first order's coefficients (and
// with DERIVS their derivatives), for the pair terms across the boundary with the previous
// lane's last order, stages each order's coefficients, and keeps their largest |Re|, |Im|.
template <bool DERIVS, class Stage>
struct WaveEfficiencySink : EfficiencySink<DERIVS> {
    using Base = EfficiencySink<DERIVS>;
    Stage stage;
    int k_first = 0;
    cx a_first = {0.0, 0.0}, b_first = {0.0, 0.0};
    cx da_first[2] = {{0.0, 0.0}, {0.0, 0.0}}, db_first[2] = {{0.0, 0.0}, {0.0, 0.0}};
    double amax = 0.0;
    __device__ void order(int k, cx a, cx b) {
        if (k == k_first) {
            a_first = a;
            b_first = b;
        }
        stage(k, a, b);
        amax = fmax(amax, fmax(fmax(fabs(a.re), fabs(a.im)), fmax(fabs(b.re), fabs(b.im))));
        Base::order(k, a, b);
    }
    // d/dx and d/dm of a, b (d/d Im m = i d/dm).
    __device__ void order_d(int k, cx a, cx b, cx da_dx, cx da_dm, cx db_dx, cx db_dm) {
        if (k == k_first) {
            a_first = a;
            b_first = b;
            da_first[0] = da_dx;
            da_first[1] = da_dm;
            db_first[0] = db_dx;
            db_first[1] = db_dm;
        }
        Base::order_d(k, a, b, da_dx, da_dm, db_dx, db_dm);
    }
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
    return v;
}

// The chains of one sphere on the whole wavefront (every lane calls it with the same sphere):
// u_k(m x) and u_k(x) into the wavefront's columns uz (n + 2 complex) and ux (n + 2 doubles),
// indexed by k, and this lane's range of orders with the recurrences' state on entry to it.
__device__ __forceinline__ OrderStart wave_prepare(double x, cx m, int n, cx *uz, double *ux,
                                                   int lane) {
    const int n_s = upward_orders(x, n);
    // u_k(x) above the upward orders (x_chain's range), kept in ux.
    if (n_s < n) {
        const int top = (int)start_index(n + 1, fabs(x));
        const int stop = n_s + 1 > 1 ? n_s + 1 : 1;
        const double inv_x = 1.0 / x;
        chain_down<double>(
            top, stop, n + 1, [&](int k) { return inv_x * (double)(2 * k + 1); },
            [&](int k, double u) { return step_real(k, inv_x, u); },
            [&](int k, double u) { ux[k] = u; }, lane);
    }
    // u_k(m x) for k = 1..n+1 (ratio_chain_ckpt's chain).
    const cx z = m * x;
    if (m.im == 0.0) {
        const double inv_rho = 1.0 / z.re;
        const int top = (int)start_index(n + 1, fabs(z.re));
        chain_down<double>(
            top, 1, n + 1, [&](int k) { return inv_rho * (double)(2 * k + 1); },
            [&](int k, double u) { return step_real(k, inv_rho, u); },
            [&](int k, double u) { uz[k] = real(u); }, lane);
    } else {
        const cx inv_z = inv(z);
        const int top = (int)start_index(n + 1, cabs(z));
        chain_down<cx>(
            top, 1, n + 1, [&](int k) { return inv_z * (double)(2 * k + 1); },
            [&](int k, cx u) { return step_complex(k, inv_z, u); },
            [&](int k, cx u) { uz[k] = u; }, lane);
    }
    __syncthreads();

    // This lane's orders k0..k1.
    int lo, hi;
    lane_chunk(n, lane, &lo, &hi);
    const int k0 = lo + 1, k1 = hi;
    const double inv_x = 1.0 / x;
    double sin_x, cos_x;
    sincos(x, &sin_x, &cos_x);
    // psi, chi below floor(x): the linear maps of this lane's orders k0..min(k1, n_s).
    Mat2<double> A = mat_identity<double>();
    for (int k = k0; k <= min(k1, n_s); ++k) {
        const double c = (double)(2 * k - 1) * inv_x;
        A = Mat2<double>{c * A.a - A.c, c * A.b - A.d, A.a, A.b};
    }
    Mat2<double> A_all;
    const Mat2<double> EA = scan_maps<false>(A, lane, &A_all);
    OrderStart st;
    st.k0 = k0;
    st.k1 = k1;
    st.psi = EA.a * sin_x + EA.b * cos_x;
    st.psi_prev = EA.c * sin_x + EA.d * cos_x;
    st.chi = EA.a * cos_x - EA.b * sin_x;
    st.chi_prev = EA.c * cos_x - EA.d * sin_x;
    // The ratio region's state at n_s (mie_orders' values at order n_s, or its start for
    // n_s = 0).
    cx t0, ixi0, r0;
    if (n_s == 0) {
        t0 = real(sin_x) * inv(mk(sin_x, -cos_x));
        ixi0 = mk(sin_x, cos_x);
        r0 = mk(0.0, 1.0);
    } else {
        const double ps = A_all.a * sin_x + A_all.b * cos_x;
        const double pp = A_all.c * sin_x + A_all.d * cos_x;
        const double ch = A_all.a * cos_x - A_all.b * sin_x;
        const double cp = A_all.c * cos_x - A_all.d * sin_x;
        const cx xi = mk(ps, -ch), xi_prev = mk(pp, -cp);
        ixi0 = inv(xi);
        t0 = real(ps) * ixi0;
        r0 = xi_prev * ixi0;
    }
    // r above floor(x): this lane's Moebius maps of orders max(k0, n_s + 1)..k1.
    const int kr0 = max(k0, n_s + 1);
    Mat2<double> R = mat_identity<double>();
    for (int k = kr0; k <= k1; ++k) {
        const double c = (double)(2 * k - 1) * inv_x;
        R = Mat2<double>{R.c, R.d, c * R.c - R.a, c * R.d - R.b};
        if (((k - kr0) & 3) == 3) R = mat_normalized(R);
    }
    R = mat_normalized(R);
    const Mat2<double> ER = scan_maps<true>(R, lane, static_cast<Mat2<double> *>(nullptr));
    // r of order kr0 - 1: ER applied to (r0, 1).
    const cx r_in = (ER.a * r0 + ER.b) * inv(ER.c * r0 + ER.d);
    // The running products of t and 1/xi over this lane's ratio orders.
    cx pt = real(1.0), px = real(1.0);
    {
        cx r = r_in;
        for (int k = kr0; k <= k1; ++k) {
            r = inv((double)(2 * k - 1) * inv_x - r);
            pt = (pt * r) * ux[k];
            px = px * r;
        }
    }
    const cx et = scan_products(pt, lane), ex = scan_products(px, lane);
    st.t = t0 * et;
    st.inv_xi = ixi0 * ex;
    st.r = kr0 == k0 ? r_in : r0;
    return st;
}

// One sphere's efficiencies on the whole wavefront (every lane calls it with the same sphere;
// lane 0's q holds the result), with DERIVS also their Jacobian jac15 (sphere_efficiencies'
// layout). uz (n + 2 complex) and ux (n + 2 doubles) are the wavefront's columns, the u_k(m x)
// and u_k(x) chains indexed by k. Each order's coefficients also go to `stage` (values only);
// *amax receives the largest |Re|, |Im| of them and *back the backscatter sum
// sum (2n+1) (-1)^n (a_n - b_n) (every lane).
template <bool DERIVS = false, class Stage = NoStage>
__device__ __forceinline__ void sphere_efficiencies_wave(double x, cx m, int n, cx *uz,
                                                         double *ux, int lane, double q[5],
                                                         Stage stage = Stage{},
                                                         double *amax = nullptr,
                                                         cx *back = nullptr,
                                                         double *jac15 = nullptr) {
    const OrderStart st = wave_prepare(x, m, n, uz, ux, lane);
    const int k0 = st.k0, k1 = st.k1;

    WaveEfficiencySink<DERIVS, Stage> s;
    s.stage = stage;
    s.k_first = k0;
    s.rk = k0 > 1 ? recip((double)k0) : 1.0;
    NoOuter none;
    if (k0 <= k1) mie_orders<DERIVS>(x, m, n, ZColumn{uz, 1}, ux, 1, none, s, &st);
    // The pair terms across the boundary with the next lane's first order k1 + 1.
    const int next = (lane + 1) & 63;
    const cx an = shfl_v(s.a_first, next), bn = shfl_v(s.b_first, next);
    const bool joined = k0 <= k1 && k1 < n;
    const double kf = (double)k1;
    const double pair = joined ? (kf * (kf + 2.0)) * recip(kf + 1.0) : 0.0;
    double g = s.g;
    if (joined) g += pair * (re_mul_conj(s.a_prev, an) + re_mul_conj(s.b_prev, bn));
    EfficiencySink<DERIVS> tot;
    tot.sca = wave_sum(s.sca);
    tot.g = wave_sum(g);
    tot.abs_sum = wave_sum(s.abs_sum);
    tot.back = cx{wave_sum(s.back.re), wave_sum(s.back.im)};
    tot.values(x, q);
    if constexpr (DERIVS) {
        const cx dan[3] = {shfl_v(s.da_first[0], next), shfl_v(s.da_first[1], next),
                           times_i(shfl_v(s.da_first[1], next))};
        const cx dbn[3] = {shfl_v(s.db_first[0], next), shfl_v(s.db_first[1], next),
                           times_i(shfl_v(s.db_first[1], next))};
#pragma unroll
        for (int p = 0; p < 3; ++p) {
            double dg = s.ds.d_g[p];
            if (joined)
                dg += pair * (re_mul_conj(s.ds.da_prev[p], an) + re_mul_conj(s.a_prev, dan[p]) +
                              re_mul_conj(s.ds.db_prev[p], bn) + re_mul_conj(s.b_prev, dbn[p]));
            tot.ds.d_g[p] = wave_sum(dg);
            tot.ds.d_sca[p] = wave_sum(s.ds.d_sca[p]);
            tot.ds.jac_abs[p] = wave_sum(s.ds.jac_abs[p]);
            tot.ds.d_back[p] = cx{wave_sum(s.ds.d_back[p].re), wave_sum(s.ds.d_back[p].im)};
        }
        tot.jacobian(x, q, true, jac15);
    }
    if (amax) *amax = wave_max(s.amax);
    if (back) *back = tot.back;
}

// The sink of one lane's orders in tangent mode (the ensembles' node solves along a direction
// of the sphere, dir 0: Re m, 1: Im m): each order's coefficients moved to c + step dc/d(dir)
// (d/d Im m = i d/dm) are staged and summed, and the absorption sum moves by step times its
// derivative along the direction, as the one-node-per-lane solve's TangentStagingSink forms
// them; the first order's moved coefficients are kept for the boundary's pair term.
template <class Stage>
struct WaveTangentSink : EfficiencySink<false> {
    using Base = EfficiencySink<false>;
    Stage stage;
    int dir = 0;
    double step = 0.0;
    int k_first = 0;
    cx a_first = {0.0, 0.0}, b_first = {0.0, 0.0};
    double amax = 0.0;
    __device__ void order_d(int k, cx a, cx b, cx, cx dam, cx, cx dbm) {
        const cx ta = dir == 0 ? dam : times_i(dam);
        const cx tb = dir == 0 ? dbm : times_i(dbm);
        const cx pa = a + ta * step, pb = b + tb * step;
        if (k == k_first) {
            a_first = pa;
            b_first = pb;
        }
        stage(k, pa, pb);
        amax = fmax(amax, fmax(fmax(fabs(pa.re), fabs(pa.im)), fmax(fabs(pb.re), fabs(pb.im))));
        Base::order(k, pa, pb);
    }
    template <class O>
    __device__ void stored(int, cx, cx, cx, cx, cx, cx, const O &) {}
    __device__ void done(double s, double j0, double j1, double j2) {
        Base::done(s + step * (dir == 0 ? j1 : j2), j0, j1, j2);
    }
};

// One sphere in tangent mode on the whole wavefront (WaveTangentSink): q the efficiencies of the
// moved coefficients (lane 0), *amax their largest |Re|, |Im| (every lane).
template <class Stage>
__device__ __forceinline__ void sphere_tangent_wave(double x, cx m, int n, cx *uz, double *ux,
                                                    int lane, int dir, double step, Stage stage,
                                                    double q[5], double *amax) {
    const OrderStart st = wave_prepare(x, m, n, uz, ux, lane);
    const int k0 = st.k0, k1 = st.k1;
    WaveTangentSink<Stage> s;
    s.stage = stage;
    s.dir = dir;
    s.step = step;
    s.k_first = k0;
    s.rk = k0 > 1 ? recip((double)k0) : 1.0;
    NoOuter none;
    if (k0 <= k1) mie_orders<true>(x, m, n, ZColumn{uz, 1}, ux, 1, none, s, &st);
    const int next = (lane + 1) & 63;
    const cx an = shfl_v(s.a_first, next), bn = shfl_v(s.b_first, next);
    double g = s.g;
    if (k0 <= k1 && k1 < n) {
        const double kf = (double)k1;
        g += (kf * (kf + 2.0)) * recip(kf + 1.0) *
             (re_mul_conj(s.a_prev, an) + re_mul_conj(s.b_prev, bn));
    }
    EfficiencySink<false> tot;
    tot.sca = wave_sum(s.sca);
    tot.g = wave_sum(g);
    tot.abs_sum = wave_sum(s.abs_sum);
    tot.back = cx{wave_sum(s.back.re), wave_sum(s.back.im)};
    tot.values(x, q);
    *amax = wave_max(s.amax);
}

// Bytes of one wavefront's columns (u_k(m x) complex, then u_k(x)) for spheres of up to nmax
// terms.
inline int64_t mie_wave_bytes(int64_t nmax) {
    const int64_t rows = nmax + 2;
    return (rows * (int64_t)(sizeof(cx) + sizeof(double)) + 255) / 256 * 256;
}

// Whether a launch of `terms` terms, the largest sphere of nmax, is solved one sphere per
// wavefront: (f = SCATRS_TUNE_SPHERE_WAVE >= 2) when its terms times f are fewer than the GPU's
// lanes (64 per wavefront, 8 wavefronts per CU) times nmax, so that one sphere per lane would
// leave most lanes idle while the largest sphere's serial recurrences run; f = 1 always, 0
// never.
inline bool wave_rule(int f, int cus, int64_t terms, int64_t nmax) {
    if (f == 0) return false;
    return f == 1 || (double)terms * f < (double)cus * 512.0 * (double)nmax;
}

}  // namespace scatrs
