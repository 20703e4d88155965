// Device kernels and host helpers of the fixed-grid ensemble path, shared by the forward entry
// (ensemble.hip) and its reverse mode (ensemble_vjp.hip). Each translation unit carries its own
// copy of the kernels (no relocatable device code), so the non-template ones are static.
//
// Kernels, per chunk of ensembles sized to the coefficient staging buffer:
//   ens_prep     - inputs checked per node (distribution, wavelength, index, validated domain),
//                  N per node, status;
//   ens_offsets  - per ensemble, the exclusive scan of N over its nodes (staging offsets);
//   ens_nodes_persist / ens_nodes_ranked - one node per lane over the chunk's nodes ranked by
//                  N, largest first (spheres: one persistent launch; layered spheres: launches
//                  whose wavefront slabs are sized by the launch's largest N): ratio chains in
//                  the slab, a_n, b_n to the staging buffer, the node's weighted scalars;
//   ens_angular  - one workgroup per ensemble: S1, S2 at +-mu from the staged coefficients and
//                  the w*pi_n, w*tau_n table, F11, F12, F33, F34 weighted and summed over the
//                  nodes in registers (fixed node order: the sums do not depend on scheduling);
//   ens_finish   - number averages, the scalar outputs, the phase matrix and the factor of the
//                  Greek coefficients;
//   ens_greek_mfma - the Greek projection with the Wigner d table as FP64 matrix-core GEMMs over
//                  the chunk's ensembles.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/scatrs_hip.h"
#include "checks.h"
#include "context.h"
#include "ensemble_common.h"
#include "layered_device.h"
#include "mie_device.h"
#include "mie_wave.h"
#include "quadrature.h"
#include "sort.h"
#include "util.h"

namespace scatrs {
// Per-node output of the node kernel: the weighted per-particle quantities
// [1, C_ext, C_sca, C_sca g, C_abs, C_back, r^2, r^3] (ensemble/engine.rs:1046-1056) times the
// quadrature weight, and the node's table factor W pi r^2 4 / x^2.
constexpr int kNodeVals = 8;

struct EnsArgs {
    const int32_t *kind;
    const double *p0, *p1, *wl, *fr, *m;
    int layers;
    int64_t n_radius;
    const double *rnodes, *rweights;  // GL rule on [-1, 1]
    const double *p2;                 // third parameter (may be null: 0)
    double2 *geom;                    // per node (r, W), written by ens_prep
    const double2 *user_rw;           // the caller's (r, W) per node, or null (from kind, p)
    const double *user_dlnw;          // with user_rw, the caller's d ln W / d p[3] per node
    // Tangent mode (scatrs_ensemble_nodes_tangent): the particle direction whose derivative is
    // taken (-1: none) and the sign of the step (TangentStagingSink).
    int tan_dir = -1;
    double tan_sign = 0.0;
};

// Ensemble e's distribution parameters.
__device__ __forceinline__ void ens_params(const EnsArgs &a, int64_t e, double p[3]) {
    p[0] = a.p0[e];
    p[1] = a.p1[e];
    p[2] = a.p2 ? a.p2[e] : 0.0;
}

struct NodeGeom {
    double r, x, W;
};

__device__ __forceinline__ NodeGeom node_geometry(const EnsArgs &a, int64_t e, int64_t i) {
    const int kind = a.kind[e];
    double p[3];
    ens_params(a, e, p);
    double t_lo, t_hi;
    radius_range(kind, p, &t_lo, &t_hi);
    const double half = 0.5 * (t_hi - t_lo), mid = 0.5 * (t_hi + t_lo);
    const double r = exp(mid + half * a.rnodes[i]);
    const double u = exp(log_density(kind, p, r, nullptr));
    NodeGeom g;
    g.r = r;
    g.x = 2.0 * kPi * r / a.wl[e];
    g.W = node_weight(kind, a.rweights[i], half, u, r);
    return g;
}

// Node i of ensemble e as ens_prep stored it: the solve kernels read it rather than evaluate
// the distributions, whose code would otherwise add to their register pressure.
__device__ __forceinline__ NodeGeom node_stored(const EnsArgs &a, int64_t e, int64_t i) {
    const double2 rw = a.geom[e * a.n_radius + i];
    NodeGeom g;
    g.r = rw.x;
    g.x = 2.0 * kPi * rw.x / a.wl[e];
    g.W = rw.y;
    return g;
}

// Status codes of the ensemble prep, beyond kValid/kInvalid/kOutside: which input failed.
enum : int { kBadDist = 1 << 4, kBadWl = 2 << 4, kBadM = 3 << 4, kBadFr = 4 << 4,
             kBadNode = 5 << 4 };

// Checks of one layered node (layered/mod.rs:229-319, layered/domain.rs:53-95), x_l = f_l x.
static __device__ int layered_code(double x, const double *fr, const double *m, int L, int allow,
                            long long *n_out) {
    *n_out = 0;
    double xl_prev = 0.0;
    for (int l = 0; l < L; ++l) {
        const double xl = (l < L - 1) ? fr[l] * x : x;
        if (!isfinite(xl)) return kInvalid;
        if (l > 0 && xl < xl_prev) return kInvalid;
        xl_prev = xl;
        const double mr = m[2 * l], mi = m[2 * l + 1];
        if (!(isfinite(mr) && isfinite(mi))) return kInvalid;
        if (mr < 0.0 || mi < 0.0) return kInvalid;
        const double am = hypot(mr, mi);
        if (!(am >= kMinIndex && am <= kMaxIndex)) return kInvalid;
    }
    const double x0 = L > 1 ? fr[0] * x : x;
    if (x0 < kMinSize || x > kMaxSize) return kInvalid;
    const long long n = default_n_max(x);
    if (n > kMaxOrder) return kInvalid;
    if (!allow) {
        bool in = x >= 1e-6 && x <= 1e4 && L <= 16;
        for (int l = 0; l < L; ++l) {
            const double mr = m[2 * l], mi = m[2 * l + 1], am = hypot(mr, mi);
            in = in && am * x <= 3e4 && mr > 0.0 && mr <= 10.0 && mi >= 0.0 && mi <= 10.0 &&
                 am >= 0.1;
        }
        if (!in) return kOutside;
    }
    *n_out = n;
    return kValid;
}

static __global__ void __launch_bounds__(256) ens_prep(EnsArgs a, int64_t count, int allow,
                                                int *node_n, Partial *partial) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t total = count * a.n_radius;
    int n = 0;
    bool ok = false;
    unsigned long long bad = ~0ull, outside = ~0ull;
    if (g < total) {
        const int64_t e = g / a.n_radius, i = g % a.n_radius;
        int code = kValid;
        if (!a.user_rw) {
            double pe[3];
            ens_params(a, e, pe);
            if (!dist_valid(a.kind[e], pe)) code = kInvalid | kBadDist;
        }
        if (code != kValid) {
        } else if (!(isfinite(a.wl[e]) && a.wl[e] > 0.0)) code = kInvalid | kBadWl;
        else if (a.layers > 1) {
            const double *fr = a.fr + e * (a.layers - 1);
            for (int l = 0; l < a.layers - 1; ++l)
                if (!(fr[l] > 0.0 && fr[l] <= 1.0) || (l > 0 && fr[l] < fr[l - 1]))
                    code = kInvalid | kBadFr;
        }
        if (code == kValid && a.user_rw) {
            const double2 rw = a.user_rw[g];
            if (!(isfinite(rw.x) && rw.x > 0.0 && isfinite(rw.y) && rw.y >= 0.0))
                code = kInvalid | kBadNode;
        }
        if (code == kValid) {
            NodeGeom geo;
            if (a.user_rw) {
                geo.r = a.user_rw[g].x;
                geo.W = a.user_rw[g].y;
                geo.x = 2.0 * kPi * geo.r / a.wl[e];
            } else {
                geo = node_geometry(a, e, i);
            }
            a.geom[g] = make_double2(geo.r, geo.W);
            long long nn = 0;
            const int c = a.layers == 1
                              ? sphere_code(geo.x, a.m[2 * e], a.m[2 * e + 1], 0, allow, &nn)
                              : layered_code(geo.x, a.fr + e * (a.layers - 1),
                                             a.m + 2 * e * a.layers, a.layers, allow, &nn);
            code = c;
            n = (int)nn;
        }
        if ((code & 15) == kInvalid) bad = (unsigned long long)g;
        if ((code & 15) == kOutside) outside = (unsigned long long)g;
        ok = code == kValid;
        node_n[g] = ok ? n : 0;
    }
    block_partial(n, ok, bad, outside, partial);
}

// The replay guard of a captured call: every node keeps its recorded N (the launches were
// planned for it), and *guard is set if the inputs changed any node's N.
static __global__ void __launch_bounds__(256) plan_guard(int *node_n, const int *node_n_rec,
                                                         int64_t total, int *guard) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g < total && node_n[g] != node_n_rec[g]) {
        node_n[g] = node_n_rec[g];
        *guard = 1;
    }
}

// Exclusive scan of N over each ensemble's nodes: node_off (terms) and ens_terms.
static __global__ void __launch_bounds__(256) ens_offsets(const int *node_n, int64_t n_radius,
                                                   int64_t *node_off, int64_t *ens_terms) {
    __shared__ int64_t part[256];
    const int64_t e = blockIdx.x;
    const int t = threadIdx.x;
    const int64_t per = (n_radius + 255) / 256;
    const int64_t lo = t * per < n_radius ? t * per : n_radius;
    const int64_t hi = lo + per < n_radius ? lo + per : n_radius;
    const int *nn = node_n + e * n_radius;
    int64_t local = 0;
    for (int64_t i = lo; i < hi; ++i) local += nn[i];
    part[t] = local;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const int64_t add = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int64_t run = part[t] - local;
    for (int64_t i = lo; i < hi; ++i) {
        node_off[e * n_radius + i] = run;
        run += nn[i];
    }
    if (t == 255) ens_terms[e] = part[255];
}

// w pi_n(mu_j), w tau_n(mu_j) at the non-negative cosines, n = 1..n_max (amplitudes.rs:52-65),
// table[(n - 1) * h + jj].
static __global__ void pt_table(const double *mu_pos, int h, int n_max, double2 *table) {
    const int jj = blockIdx.x * blockDim.x + threadIdx.x;
    if (jj >= h) return;
    const double mu = mu_pos[jj];
    double pi_prev = 0.0, pi = 1.0;
    for (int n = 1; n <= n_max; ++n) {
        const double nf = n, n1 = nf + 1.0;
        const double tau = nf * mu * pi - n1 * pi_prev;
        const double w = (double)(2 * n + 1) / (nf * n1);
        table[(size_t)(n - 1) * h + jj] = make_double2(w * pi, w * tau);
        const double pi_next = ((double)(2 * n + 1) * mu * pi - n1 * pi_prev) / nf;
        pi_prev = pi;
        pi = pi_next;
    }
}

// Rows of a sphere slab's checkpoint column: one ratio per kZBlock orders.
__host__ __device__ inline int64_t sphere_ckpt_rows(int64_t rows) { return (rows - 2) / kZBlock + 2; }

// Complex elements of one wavefront's scratch slab for nodes of at most rows - 2 terms, 64
// lanes interleaved: spheres keep the ratio checkpoints and a real column of u_k(x); layered
// spheres 2 L - 1 complex columns and a real one.
__host__ __device__ inline int64_t slab_elems(int layers, int64_t rows) {
    if (layers == 1) return 64 * (sphere_ckpt_rows(rows) + (rows + 1) / 2);
    return rows * 64 * 2 * layers;
}

// Writes a node's coefficients to the staging buffer, interleaved: in FP64 (a_1, b_1, a_2, b_2,
// ...: one 32-byte stream per lane, whose cache lines fill every four orders) or, for the
// fp16-split angular products, in fp32 (one float4 (a.re, a.im, b.re, b.im) per order); and
// accumulates its efficiency sums. MODE (fixed per launch, so that the order loop carries no
// branch on it): kStageF64; kStageF32 (one float4 per order straight to the buffer); kStageTile
// (spheres' fp32 form: each order's float4 goes to the lane's slot of the order loop's LDS tile,
// its u_{k+1}(mx) consumed by then, and the block's orders go out together at its end, 16 B x 16
// contiguous per lane, so that the staging buffer receives whole cache lines).
constexpr int kStageF64 = 0, kStageF32 = 1, kStageTile = 2;

template <int MODE>
struct StagingSink {
    cx *ab_out;
    float4 *ab32;
    float4 *tile32 = nullptr;
    EfficiencySink<false> eff;
    double amax = 0.0;  // the largest |Re|, |Im| of the node's coefficients
    __device__ void order(int k, cx a, cx b) {
        if constexpr (MODE == kStageTile) {
            tile32[((k - 1) % kZBlock) * 64] =
                make_float4((float)a.re, (float)a.im, (float)b.re, (float)b.im);
        } else if constexpr (MODE == kStageF32) {
            ab32[k - 1] = make_float4((float)a.re, (float)a.im, (float)b.re, (float)b.im);
        } else {
            ab_out[2 * (k - 1)] = a;
            ab_out[2 * (k - 1) + 1] = b;
        }
        amax = fmax(amax, fmax(fmax(fabs(a.re), fabs(a.im)), fmax(fabs(b.re), fabs(b.im))));
        eff.order(k, a, b);
    }
    __device__ void block_end(int k0, int kend) {
        if constexpr (MODE == kStageTile) {
            for (int k = k0; k <= kend; ++k) ab32[k - 1] = tile32[(k - k0) * 64];
        } else {
            (void)k0, (void)kend;
        }
    }
    __device__ void order_d(int, cx, cx, cx, cx, cx, cx) {}
    template <class O>
    __device__ void stored(int, cx, cx, cx, cx, cx, cx, const O &) {}
    __device__ void done(double s, double j0, double j1, double j2) { eff.done(s, j0, j1, j2); }
    __device__ void zero_absorption() { eff.zero_absorption(); }
};

// Tangent mode: a node's coefficients moved along one of the particle's directions,
// c' = c + sign eps dc/d(theta) with eps = 1 / (1 + x), staged as StagingSink stages c (FP64 or
// fp32), with the efficiency sums of c'. Every output of an ensemble is linear or quadratic in
// its nodes' coefficients (S1, S2 linear; F, C_sca, C_sca g, C_back quadratic), so
// [q(c + eps dc) - q(c - eps dc)] / (2 eps) is dq exactly (to rounding): two calls, sign +1
// and -1, with the node weights scaled by 1 / (2 eps), give the derivative of the weighted
// sums as their difference. The absorption sum (C_ext = C_sca + C_abs, efficiencies.rs:169-187)
// moves by sign eps times its derivative along the direction, from the absorption identity's
// Jacobian: the outer layer's from the solve's (done), an inner layer's summed order by order
// from its sensitivities (stored), as the reverse mode's AdjointSink forms them. Directions as
// the reference's (ensemble/particle.rs:221-249): 2 l, 2 l + 1 are Re m, Im m of layer l
// (d/d Im m = i d/dm), 2 L + j the radius fraction j (x_L d/dx_j at a fixed x_L).
template <int L>
struct TangentStagingSink {
    cx *ab_out;
    float4 *ab32;  // not null: the fp32 form
    EfficiencySink<false> eff;
    double amax = 0.0;
    int dir;
    double step;  // sign eps
    double x;     // the outer size parameter (radius-fraction directions)
    double dabs = 0.0;  // the absorption sum's derivative along the direction (inner layers:
                        // summed over the orders; after done: the direction's)
    cx pa, pb;    // the order's moved coefficients before the inner layers' part
    __device__ void put(int k, cx a, cx b) {
        if (ab32) {
            ab32[k - 1] = make_float4((float)a.re, (float)a.im, (float)b.re, (float)b.im);
        } else {
            ab_out[2 * (k - 1)] = a;
            ab_out[2 * (k - 1) + 1] = b;
        }
        amax = fmax(amax, fmax(fmax(fabs(a.re), fabs(a.im)), fmax(fabs(b.re), fabs(b.im))));
        eff.order(k, a, b);
    }
    __device__ void order(int k, cx a, cx b) { put(k, a, b); }
    __device__ void order_d(int k, cx a, cx b, cx, cx dam, cx, cx dbm) {
        cx ta = real(0.0), tb = real(0.0);
        if (dir == 2 * (L - 1)) {
            ta = dam;
            tb = dbm;
        } else if (dir == 2 * (L - 1) + 1) {
            ta = times_i(dam);
            tb = times_i(dbm);
        }
        pa = a + ta * step;
        pb = b + tb * step;
        if constexpr (L == 1) put(k, pa, pb);
    }
    template <class O>
    __device__ void stored(int k, cx inv_da, cx inv_db, cx g_a, cx g_b, cx d_m, cx md,
                           const O &o) {
        if constexpr (L > 1) {
            // the inner layer (or interface) the direction moves: l < L - 1
            const int l = dir < 2 * L ? dir / 2 : dir - 2 * L;
            if (l < L - 1) {
                const cx mL = o.m[L - 1];
                const cx inv_mL = inv(mL);
                const bool frac = dir >= 2 * L;
                const int c = frac ? 0 : 1;  // dh columns: [d/dx_l, d/dm_l] for a, then for b
                cx ba = o.dh[l][c] * inv_mL, bb = mL * o.dh[l][2 + c];
                if (!frac && (dir & 1)) {  // d/d Im m = i d/dm
                    ba = times_i(ba);
                    bb = times_i(bb);
                }
                cx da = -(times_i(ba) * (inv_da * inv_da));
                cx db = -(times_i(bb) * (inv_db * inv_db));
                double ab = (double)(2 * k + 1) *
                            (absorption_term(d_m, ba, ba * g_a, norm2(inv_da)) +
                             absorption_term(md, bb, bb * g_b, norm2(inv_db)));
                if (frac) {
                    da = da * x;
                    db = db * x;
                    ab *= x;
                }
                dabs += ab;
                pa = pa + da * step;
                pb = pb + db * step;
            }
            put(k, pa, pb);
        } else {
            (void)k, (void)inv_da, (void)inv_db, (void)g_a, (void)g_b, (void)d_m, (void)md,
                (void)o;
        }
    }
    // The absorption sum's derivative along the direction: the outer layer's from the solve.
    __device__ double absorption_derivative(double j1, double j2) const {
        if (dir == 2 * (L - 1)) return j1;
        if (dir == 2 * (L - 1) + 1) return j2;
        return dabs;
    }
    __device__ void done(double s, double j0, double j1, double j2) {
        dabs = absorption_derivative(j1, j2);
        eff.done(s + step * dabs, j0, j1, j2);
    }
    // A lossless layered sphere absorbs nothing and stays lossless as any x_l or Re m_l changes
    // (layered/kernel.rs:453-464); along an Im m direction its absorption moves from 0.
    __device__ void zero_absorption() {
        eff.zero_absorption();
        if (dir < 2 * L && (dir & 1)) eff.abs_sum = step * dabs;
    }
};

// The exponent s that brings a node's largest coefficient component amax into [2^9, 2^10) by
// 2^s (|a_n|, |b_n| <= 1), within [0, 100]: the fp16-split angular products scale the node's
// coefficients by 2^s, so that their slices keep fp16's full precision (those of small
// particles, |a_1| ~ x^3, would otherwise fall below its normal range), and its table factor by
// 2^-2s, both exactly.
constexpr int kCoefficientBits = 10;

__device__ __forceinline__ int coefficient_exponent(double amax) {
    if (!(amax > 0.0)) return kCoefficientBits - 1;
    const int s = -ilogb(amax) - 1 + kCoefficientBits;
    return s < 0 ? 0 : (s > 100 ? 100 : s);
}

// One node's forward solve: node g of the chunk is ensemble e0 + g / n_radius, radius node
// g % n_radius. Writes its coefficients to the staging buffer, its weighted per-particle
// quantities to vals, its table factor to tabc, its coefficient exponent to xexp and, if
// back_out is not null, its backscatter sum sum (2n+1) (-1)^n (a_n - b_n). slab is the lane's
// column in its wavefront's scratch slab of `rows` rows (slab_elems); tile the lane's column of
// the wavefront's LDS tile (spheres). With stage32 the staging buffer holds the fp32 form.
template <int L>
__device__ __forceinline__ void forward_node(const EnsArgs &a, int64_t e0, int64_t g,
                                             const int *node_n, const int64_t *node_off,
                                             const int64_t *ens_base, cx *slab, int64_t rows,
                                             cx *tile, cx *staging, int stage32, double *vals,
                                             double *tabc, int *xexp, cx *back_out) {
    const int64_t e = e0 + g / a.n_radius, i = g % a.n_radius;
    const int64_t gi = e * a.n_radius + i;
    const NodeGeom geo = node_stored(a, e, i);
    const int n = node_n[gi];
    const int64_t off = ens_base[e] + node_off[gi];
    auto solve = [&](auto &sink) __attribute__((always_inline)) {
        using Sink = std::remove_reference_t<decltype(sink)>;
        sink.ab_out = staging + 2 * off;
        sink.ab32 = reinterpret_cast<float4 *>(staging) + off;
        sink.tile32 = reinterpret_cast<float4 *>(tile);
        if constexpr (L == 1) {
            const cx m = mk(a.m[2 * e], a.m[2 * e + 1]);
            // Checkpoints in the first column, u_k(x) in the second, the block being used in
            // the wavefront's LDS tile. The node first runs whole on the plain path (branch-free
            // blocks, no per-block state for a redo) and, in the rare lane whose inverses left
            // the plain range, again on the checked path, restaging its coefficients.
            double *ux = real_column(slab, sphere_ckpt_rows(rows) * 64);
            x_chain(geo.x, n, upward_orders(geo.x, n), ux, 64);
            NoOuter none;
            const Sink fresh = sink;
            bool ok;
            {
                const auto zs = ratio_chain_ckpt<kZBlock, true>(m, geo.x, n, slab, 64, tile);
                mie_orders<false>(geo.x, m, n, zs, ux, 64, none, sink);
                ok = zs.plain_ok();
            }
            if (!ok) {
                sink = fresh;
                const ZCheckpoint zc = ratio_chain_ckpt(m, geo.x, n, slab, 64, tile);
                mie_orders<false>(geo.x, m, n, zc, ux, 64, none, sink);
            }
        } else {
            (void)tile;
            layered_solve<L, false>(geo.x, a.fr + e * (L - 1), a.m + 2 * e * L, n, slab, rows,
                                    sink);
        }
        return sink.amax;
    };
    EfficiencySink<false> eff;
    double amax;
    if (stage32) {
        StagingSink<L == 1 ? kStageTile : kStageF32> sink;
        amax = solve(sink);
        eff = sink.eff;
    } else {
        StagingSink<kStageF64> sink;
        amax = solve(sink);
        eff = sink.eff;
    }
    double q[5];
    eff.values(geo.x, q);
    const double area = kPi * geo.r * geo.r;
    const double W = geo.W;
    double *v = vals + g * kNodeVals;
    v[0] = W;
    v[1] = W * (area * q[0]);
    v[2] = W * (area * q[1]);
    v[3] = W * (area * q[1] * q[4]);
    v[4] = W * (area * q[2]);
    v[5] = W * (area * q[3]);
    v[6] = W * (geo.r * geo.r);
    v[7] = W * (geo.r * geo.r * geo.r);
    tabc[g] = W * (area * (4.0 / (geo.x * geo.x)));
    xexp[g] = coefficient_exponent(amax);
    if (back_out) back_out[g] = eff.back;
}

// forward_node in tangent mode (a.tan_dir >= 0): the solve with derivatives, the coefficients
// moved along the direction (TangentStagingSink), the node's weight times 1 / (2 eps).
template <int L>
__device__ __forceinline__ void tangent_node(const EnsArgs &a, int64_t e0, int64_t g,
                                             const int *node_n, const int64_t *node_off,
                                             const int64_t *ens_base, cx *slab, int64_t rows,
                                             cx *tile, cx *staging, int stage32, double *vals,
                                             double *tabc, int *xexp) {
    const int64_t e = e0 + g / a.n_radius, i = g % a.n_radius;
    const int64_t gi = e * a.n_radius + i;
    const NodeGeom geo = node_stored(a, e, i);
    const int n = node_n[gi];
    const int64_t off = ens_base[e] + node_off[gi];
    const double eps = 1.0 / (1.0 + geo.x);
    TangentStagingSink<L> sink;
    sink.ab_out = staging + 2 * off;
    sink.ab32 = stage32 ? reinterpret_cast<float4 *>(staging) + off : nullptr;
    sink.dir = a.tan_dir;
    sink.step = a.tan_sign * eps;
    sink.x = geo.x;
    if constexpr (L == 1) {
        const cx m = mk(a.m[2 * e], a.m[2 * e + 1]);
        double *ux = real_column(slab, sphere_ckpt_rows(rows) * 64);
        const ZCheckpoint zs = ratio_chain_ckpt(m, geo.x, n, slab, 64, tile);
        x_chain(geo.x, n, upward_orders(geo.x, n), ux, 64);
        NoOuter none;
        mie_orders<true>(geo.x, m, n, zs, ux, 64, none, sink);
    } else {
        (void)tile;
        layered_solve<L, true>(geo.x, a.fr + e * (L - 1), a.m + 2 * e * L, n, slab, rows, sink);
    }
    double q[5];
    sink.eff.values(geo.x, q);
    const double area = kPi * geo.r * geo.r;
    const double W = geo.W / (2.0 * eps);
    double *v = vals + g * kNodeVals;
    v[0] = W;
    v[1] = W * (area * q[0]);
    v[2] = W * (area * q[1]);
    v[3] = W * (area * q[1] * q[4]);
    v[4] = W * (area * q[2]);
    v[5] = W * (area * q[3]);
    v[6] = W * (geo.r * geo.r);
    v[7] = W * (geo.r * geo.r * geo.r);
    tabc[g] = W * (area * (4.0 / (geo.x * geo.x)));
    xexp[g] = coefficient_exponent(sink.amax);
}

// Spheres: persistent; each wavefront takes the next 64 ranked nodes from the counter, so the
// nodes of largest N start first and the rest fill in behind them, and the 64 lanes of a
// wavefront solve nodes of about the same N. Slabs of `rows` rows (the chunk's largest N).
static __global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void ens_nodes_persist(
    EnsArgs a, int64_t e0, int64_t chunk_nodes, const int *order, const int *node_n,
    const int64_t *node_off, const int64_t *ens_base, unsigned long long *counter, cx *scratch,
    int64_t rows, cx *staging, int stage32, double *vals, double *tabc, int *xexp,
    cx *back_out) {
    __shared__ cx ztile[4][kZBlock * 64];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    cx *tile = &ztile[threadIdx.x >> 6][lane];
    cx *slab = scratch + wave * slab_elems(1, rows) + lane;
    while (true) {
        unsigned long long chunk = 0;
        if (lane == 0) chunk = atomicAdd(counter, 1ull);
        chunk = __shfl(chunk, 0, 64);
        if ((int64_t)chunk * 64 >= chunk_nodes) break;
        const int64_t idx = (int64_t)chunk * 64 + lane;
        if (idx < chunk_nodes)
            forward_node<1>(a, e0, order[idx], node_n, node_off, ens_base, slab, rows, tile,
                            staging, stage32, vals, tabc, xexp, back_out);
    }
}

// The node solves of a small launch (spheres), one node per wavefront (blocks of one
// wavefront) from the shared counter over the ranked nodes: what forward_node writes (staging,
// values, table factor, coefficient exponent, backscatter sum), its orders split over the
// lanes (sphere_efficiencies_wave); in tangent mode (a.tan_dir >= 0) what tangent_node writes
// (sphere_tangent_wave, the weight times 1 / (2 eps)). staging == nullptr: nothing staged (a
// call without angular products).
static __global__ __launch_bounds__(64) void ens_nodes_wave(
    EnsArgs a, int64_t e0, int64_t chunk_nodes, const int *order, const int *node_n,
    const int64_t *node_off, const int64_t *ens_base, unsigned long long *counter, char *cols,
    int64_t rows, int64_t wave_bytes, cx *staging, int stage32, double *vals, double *tabc,
    int *xexp, cx *back_out) {
    const int lane = threadIdx.x;
    char *base = cols + (int64_t)blockIdx.x * wave_bytes;
    cx *uz = reinterpret_cast<cx *>(base);
    double *ux = reinterpret_cast<double *>(base + rows * (int64_t)sizeof(cx));
    while (true) {
        unsigned long long idx = 0;
        if (lane == 0) idx = atomicAdd(counter, 1ull);
        idx = __shfl(idx, 0, 64);
        if ((int64_t)idx >= chunk_nodes) break;
        const int64_t g = order[idx];
        const int64_t e = e0 + g / a.n_radius, i = g % a.n_radius;
        const NodeGeom geo = node_stored(a, e, i);
        const int n = node_n[e * a.n_radius + i];
        const cx m = mk(a.m[2 * e], a.m[2 * e + 1]);
        double q[5], amax = 0.0;
        cx back = {0.0, 0.0};
        if (a.tan_dir >= 0) {
            const double eps = 1.0 / (1.0 + geo.x);
            const double step = a.tan_sign * eps;
            const int64_t off = ens_base[e] + node_off[e * a.n_radius + i];
            if (!staging)
                sphere_tangent_wave(geo.x, m, n, uz, ux, lane, a.tan_dir, step, NoStage{}, q, &amax);
            else if (stage32)
                sphere_tangent_wave(geo.x, m, n, uz, ux, lane, a.tan_dir, step,
                                    StageF32{reinterpret_cast<float4 *>(staging) + off}, q, &amax);
            else
                sphere_tangent_wave(geo.x, m, n, uz, ux, lane, a.tan_dir, step,
                                    StageF64{staging + 2 * off}, q, &amax);
        } else if (!staging) {
            sphere_efficiencies_wave(geo.x, m, n, uz, ux, lane, q, NoStage{}, &amax, &back);
        } else {
            const int64_t off = ens_base[e] + node_off[e * a.n_radius + i];
            if (stage32)
                sphere_efficiencies_wave(geo.x, m, n, uz, ux, lane, q,
                                         StageF32{reinterpret_cast<float4 *>(staging) + off},
                                         &amax, &back);
            else
                sphere_efficiencies_wave(geo.x, m, n, uz, ux, lane, q, StageF64{staging + 2 * off},
                                         &amax, &back);
        }
        if (lane == 0) {
            const double area = kPi * geo.r * geo.r;
            const double W = a.tan_dir >= 0 ? geo.W / (2.0 * (1.0 / (1.0 + geo.x))) : geo.W;
            double *v = vals + g * kNodeVals;
            v[0] = W;
            v[1] = W * (area * q[0]);
            v[2] = W * (area * q[1]);
            v[3] = W * (area * q[1] * q[4]);
            v[4] = W * (area * q[2]);
            v[5] = W * (area * q[3]);
            v[6] = W * (geo.r * geo.r);
            v[7] = W * (geo.r * geo.r * geo.r);
            tabc[g] = W * (area * (4.0 / (geo.x * geo.x)));
            xexp[g] = coefficient_exponent(amax);
            if (back_out) back_out[g] = back;
        }
        // The columns are rewritten by the next node only after every lane has read them.
        __syncthreads();
    }
}

// Tangent mode, spheres: persistent over the ranked nodes, as ens_nodes_persist.
static __global__ __launch_bounds__(256) void ens_nodes_tangent_persist(
    EnsArgs a, int64_t e0, int64_t chunk_nodes, const int *order, const int *node_n,
    const int64_t *node_off, const int64_t *ens_base, unsigned long long *counter, cx *scratch,
    int64_t rows, cx *staging, int stage32, double *vals, double *tabc, int *xexp) {
    __shared__ cx ztile[4][kZBlock * 64];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    cx *tile = &ztile[threadIdx.x >> 6][lane];
    cx *slab = scratch + wave * slab_elems(1, rows) + lane;
    while (true) {
        unsigned long long chunk = 0;
        if (lane == 0) chunk = atomicAdd(counter, 1ull);
        chunk = __shfl(chunk, 0, 64);
        if ((int64_t)chunk * 64 >= chunk_nodes) break;
        const int64_t idx = (int64_t)chunk * 64 + lane;
        if (idx < chunk_nodes)
            tangent_node<1>(a, e0, order[idx], node_n, node_off, ens_base, slab, rows, tile,
                            staging, stage32, vals, tabc, xexp);
    }
}

// Tangent mode, layered spheres: one launch covers the ranked nodes [start, end), one group of
// 64 per wavefront and no work loop (slabs sized by the launch's first node), as
// ens_nodes_ranked.
template <int L>
__global__ __launch_bounds__(256) void ens_nodes_tangent(EnsArgs a, int64_t e0, const int *order,
                                                         int64_t start, int64_t end,
                                                         const int *node_n,
                                                         const int64_t *node_off,
                                                         const int64_t *ens_base, cx *scratch,
                                                         int64_t rows, cx *staging, int stage32,
                                                         double *vals, double *tabc, int *xexp) {
    __shared__ cx ztile[4][L == 1 ? kZBlock * 64 : 1];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    cx *slab = scratch + wave * slab_elems(L, rows) + lane;
    cx *tile = L == 1 ? &ztile[threadIdx.x >> 6][lane] : nullptr;
    const int64_t idx = start + wave * 64 + lane;
    if (idx < end)
        tangent_node<L>(a, e0, order[idx], node_n, node_off, ens_base, slab, rows, tile, staging,
                        stage32, vals, tabc, xexp);
}

// Layered spheres: one launch covers the ranked nodes [start, end), one group of 64 per
// wavefront and no work loop; `rows` (the slab's) is set by the launch's first node. The layered
// solves keep enough state per lane that their registers spill; run inside a persistent work
// loop, that solve has given wrong results, so these launches carry none.
template <int L>
__global__ __launch_bounds__(256) void ens_nodes_ranked(EnsArgs a, int64_t e0, const int *order,
                                                        int64_t start, int64_t end,
                                                        const int *node_n,
                                                        const int64_t *node_off,
                                                        const int64_t *ens_base, cx *scratch,
                                                        int64_t rows, cx *staging, int stage32,
                                                        double *vals, double *tabc, int *xexp,
                                                        cx *back_out) {
    __shared__ cx ztile[4][L == 1 ? kZBlock * 64 : 1];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    cx *slab = scratch + wave * slab_elems(L, rows) + lane;
    cx *tile = L == 1 ? &ztile[threadIdx.x >> 6][lane] : nullptr;
    const int64_t idx = start + wave * 64 + lane;
    if (idx < end)
        forward_node<L>(a, e0, order[idx], node_n, node_off, ens_base, slab, rows, tile, staging,
                        stage32, vals, tabc, xexp, back_out);
}

// Sort keys of the chunk's nodes: bucket_of(N), N descending.
static __global__ void ens_node_keys(int64_t e0, int64_t n_radius, int64_t chunk_nodes,
                                     const int *node_n, int *keys) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g < chunk_nodes) keys[g] = bucket_of(node_n[e0 * n_radius + g]);
}

constexpr int kTileNodes = 4;
constexpr int kChunkOrders = 64;

// One workgroup per ensemble; thread t owns the non-negative cosine jj = t (+ blockDim per
// pass) and its mirror. For each tile of 4 nodes the coefficients are staged in LDS 64 orders
// at a time; per order, the E/O parity split: A += G1 u, B += G2 v, C += G1 v, D += G2 u with
// G1, G2 = w pi_n, w tau_n (odd n) or w tau_n, w pi_n (even n) and u, v = a_n, b_n (odd) or
// b_n, a_n (even), so that S1(+-mu) = A +- B and S2(+-mu) = C +- D.
static __global__ __launch_bounds__(512) void ens_angular(int64_t e0, int64_t n_radius, int n_angles,
                                                   int h, const int *node_n,
                                                   const int64_t *node_off,
                                                   const int64_t *ens_base, const cx *staging,
                                                   const double *tabc, const double2 *table,
                                                   double *tab_out) {
    __shared__ cx lds_a[kTileNodes][kChunkOrders];
    __shared__ cx lds_b[kTileNodes][kChunkOrders];
    const int64_t e = e0 + blockIdx.x;
    const int t = threadIdx.x;
    const int64_t base = ens_base[e];
    const int *nn = node_n + e * n_radius;
    const int64_t *noff = node_off + e * n_radius;
    const int64_t gbase = blockIdx.x * n_radius;  // node index within the chunk
    double *out = tab_out + (int64_t)blockIdx.x * 4 * n_angles;
    for (int pass = 0; pass * (int)blockDim.x < h; ++pass) {
        const int jj = pass * blockDim.x + t;
        const bool active = jj < h;
        double T[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};  // [sign][F11, F12, F33, F34]
        for (int64_t i0 = 0; i0 < n_radius; i0 += kTileNodes) {
            int ns[kTileNodes];
            int nmax = 0;
#pragma unroll
            for (int q = 0; q < kTileNodes; ++q) {
                ns[q] = (i0 + q < n_radius) ? nn[i0 + q] : 0;
                nmax = ns[q] > nmax ? ns[q] : nmax;
            }
            cx A[kTileNodes], B[kTileNodes], C[kTileNodes], D[kTileNodes];
#pragma unroll
            for (int q = 0; q < kTileNodes; ++q) A[q] = B[q] = C[q] = D[q] = real(0.0);
            for (int kc = 0; kc < nmax; kc += kChunkOrders) {
                __syncthreads();
                for (int idx = t; idx < kTileNodes * kChunkOrders; idx += blockDim.x) {
                    const int q = idx / kChunkOrders, k = idx % kChunkOrders;
                    const int n = kc + k;
                    cx va = real(0.0), vb = real(0.0);
                    if (n < ns[q]) {
                        const cx *c = staging + 2 * (base + noff[i0 + q]);
                        va = c[2 * n];
                        vb = c[2 * n + 1];
                    }
                    lds_a[q][k] = va;
                    lds_b[q][k] = vb;
                }
                __syncthreads();
                if (active) {
                    const int kend = nmax - kc < kChunkOrders ? nmax - kc : kChunkOrders;
                    for (int k = 0; k < kend; ++k) {
                        const int n = kc + k + 1;  // the order
                        const double2 pt = table[(size_t)(n - 1) * h + jj];
                        const double g1 = (n & 1) ? pt.x : pt.y;
                        const double g2 = (n & 1) ? pt.y : pt.x;
#pragma unroll
                        for (int q = 0; q < kTileNodes; ++q) {
                            const cx a = lds_a[q][k], b = lds_b[q][k];
                            const cx u = (n & 1) ? a : b;
                            const cx v = (n & 1) ? b : a;
                            A[q] = A[q] + u * g1;
                            B[q] = B[q] + v * g2;
                            C[q] = C[q] + v * g1;
                            D[q] = D[q] + u * g2;
                        }
                    }
                }
            }
            if (active) {
#pragma unroll
                for (int q = 0; q < kTileNodes; ++q) {
                    if (i0 + q >= n_radius) continue;
                    const double c = tabc[gbase + i0 + q];
#pragma unroll
                    for (int sgn = 0; sgn < 2; ++sgn) {
                        const cx s1 = sgn == 0 ? A[q] + B[q] : A[q] - B[q];
                        const cx s2 = sgn == 0 ? C[q] + D[q] : C[q] - D[q];
                        const double n1 = norm2(s1), n2 = norm2(s2);
                        const cx p = s2 * conj(s1);
                        T[sgn][0] += c * (0.5 * (n1 + n2));
                        T[sgn][1] += c * (0.5 * (n2 - n1));
                        T[sgn][2] += c * p.re;
                        T[sgn][3] += c * p.im;
                    }
                }
            }
        }
        if (active) {
            const int jp = n_angles - h + jj;     // the cosine's index in the ascending rule
            const int jm = n_angles - 1 - jp;     // its mirror
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                out[k * n_angles + jp] = T[0][k];
                if (jm != jp) out[k * n_angles + jm] = T[1][k];
            }
        }
    }
}

// The angular products on FP64 matrix cores (v_mfma_f64_16x16x4f64), as GEMMs per group of
// 16 radius nodes: C1^T = X^T G1^T and C2^T = X^T G2^T with G1, G2 (angles x orders) the
// w pi_n, w tau_n table by parity and X (orders x 64) the nodes' u, v (a_n, b_n by parity)
// split into re, im. The coefficients are the A operand: row 4c + q of a 16-row tile is
// component c of node q, (u.re, u.im, v.re, v.im)[c]; the table is the B operand, a
// wavefront's 16 columns its cosines. v_mfma_f64_16x16x4f64 leaves D[i][j] in lane
// j + 16 (i % 4), element i / 4, so each lane receives the four components of node q = lane / 16
// of the tile at its cosine j = lane % 16, forms S1(+-mu) = A +- B, S2(+-mu) = C +- D (A, C =
// G1 u, G1 v; B, D = G2 v, G2 u) and F11, F12, F33, F34 itself, times the node's factor, summed
// over the groups in registers; the four node lanes of a cosine are added at the end, always in
// the same order. A workgroup is one ensemble and 128 non-negative cosines (8 wavefronts x 16);
// the coefficients of a group are staged in LDS 32 orders at a time.
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

constexpr int kGroupNodes = 16;
constexpr int kMfmaOrders = 32;
constexpr int kSliceAngles = 128;

// With splits > 1 (small batches, whose ensembles and cosine slices alone do not fill the
// chip) the node groups of an ensemble are dealt over `splits` blocks, split sp taking groups
// sp, sp + splits, ...; each writes its partial table to tab_out + sp * (ne 4 n_angles), and
// ens_split_reduce adds the splits in order.
// With lds_nodes the ensemble's node list sits in dynamic LDS (mfma_lds_bytes): N and the staging
// offset of each node, then each group's largest N, so that moving to the next item issues no
// dependent global loads; without it (more than kMfmaLdsNodes nodes) they come from memory.
constexpr int64_t kMfmaLdsNodes = 8192;

__host__ __device__ inline size_t mfma_lds_bytes(int64_t n_radius) {
    return (size_t)(2 * n_radius + (n_radius + kGroupNodes - 1) / kGroupNodes) * sizeof(int);
}

static __global__ __launch_bounds__(512) void ens_angular_mfma(
    int64_t e0, int64_t n_radius, int n_angles, int h, int slices, int splits, int64_t ne,
    const int *node_n, const int64_t *node_off, const int64_t *ens_base, const cx *staging,
    const double *tabc, const double2 *table, double *tab_out, int lds_nodes) {
    __shared__ cx lds_a[kGroupNodes][kMfmaOrders];
    __shared__ cx lds_b[kGroupNodes][kMfmaOrders];
    extern __shared__ int mf_node[];  // [n_radius] N, [n_radius] offsets, [groups] largest N
    const int sp = blockIdx.x % splits;
    const int64_t eb = (blockIdx.x / splits) / slices;
    const int slice = (blockIdx.x / splits) % slices;
    const int64_t e = e0 + eb;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int row = lane & 15, kq = lane >> 4;  // A: row 4c + q, order; B: cosine, order
    const int comp = row >> 2, qin = row & 3;   // the A row's component and node of the tile
    const int jj = slice * kSliceAngles + wave * 16 + row;  // this lane's B-operand cosine
    const int64_t base = ens_base[e];
    const int *nn = node_n + e * n_radius;
    const int64_t *noff = node_off + e * n_radius;
    const int64_t gbase = eb * n_radius;
    // [sign][F11, F12, F33, F34] at cosine jj of the lane's nodes (4 tt + kq of each group)
    double tp[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    // Software pipeline over the items (group g0, 32-order chunk kc) of this block: while an
    // item's products run, the next item's staged coefficients (this thread's (node, order)
    // pair) and its eight table steps load into registers, so that no load waits in the
    // product loop; the products and their order are those of the unpipelined loop.
    const int sq = t / kMfmaOrders, sk = t % kMfmaOrders;  // this thread's staging pair
    const int64_t gstep = (int64_t)splits * kGroupNodes;
    const int nr = (int)n_radius;
    if (lds_nodes) {
        const int groups = (nr + kGroupNodes - 1) / kGroupNodes;
        for (int i = t; i < groups; i += blockDim.x) mf_node[2 * nr + i] = 0;
        for (int i = t; i < nr; i += blockDim.x) {
            mf_node[i] = nn[i];
            mf_node[nr + i] = (int)noff[i];
        }
        __syncthreads();
        for (int i = t; i < nr; i += blockDim.x)
            atomicMax(&mf_node[2 * nr + i / kGroupNodes], mf_node[i]);
        __syncthreads();
    }
    auto node_nq = [&](int64_t i) { return lds_nodes ? mf_node[i] : nn[i]; };
    auto node_of = [&](int64_t i) -> int64_t { return lds_nodes ? (int64_t)mf_node[nr + i] : noff[i]; };
    auto group_nmax = [&](int64_t g0) {
        if (lds_nodes) return mf_node[2 * nr + (int)(g0 / kGroupNodes)];
        int nm = 0;
        for (int q = 0; q < kGroupNodes; ++q) {
            const int nq = g0 + q < n_radius ? nn[g0 + q] : 0;
            nm = nq > nm ? nq : nm;
        }
        return nm;
    };
    // The item's coefficient pair (zero beyond its node's N) and table steps (zero beyond the
    // group's largest N or the cosines).
    auto load = [&](int64_t g0, int kc, int nm, cx &va, cx &vb, double2 (&tb)[8])
                    __attribute__((always_inline)) {
        const int n = kc + sk;
        const int nq = g0 + sq < n_radius ? node_nq(g0 + sq) : 0;
        va = real(0.0);
        vb = real(0.0);
        if (n < nq) {
            const cx *c = staging + 2 * (base + node_of(g0 + sq));
            va = c[2 * n];
            vb = c[2 * n + 1];
        }
        // Loaded from a clamped address in every lane (the load count per item is fixed, so the
        // product loop's waits cover exactly the previous item's loads), masked after arrival.
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            const int nt = kc + 4 * st + kq + 1;
            const bool ok = jj < h && nt <= nm;
            const double2 v = table[(size_t)(ok ? nt - 1 : 0) * h + (jj < h ? jj : 0)];
            tb[st] = ok ? v : make_double2(0.0, 0.0);
        }
    };
    int64_t g0 = (int64_t)sp * kGroupNodes;
    if (g0 >= n_radius) g0 = n_radius;  // no group for this split: straight to the output
    int nmax = g0 < n_radius ? group_nmax(g0) : 0;
    int kc = 0;
    cx pa = real(0.0), pb = real(0.0);
    double2 tn[8];
    if (g0 < n_radius) load(g0, 0, nmax, pa, pb, tn);
    mfma_d4 acc1[4], acc2[4];
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) {
        acc1[tt] = mfma_d4{0, 0, 0, 0};
        acc2[tt] = mfma_d4{0, 0, 0, 0};
    }
    while (g0 < n_radius) {
        __syncthreads();  // the previous item's products are done with the LDS tile
        if (t < kGroupNodes * kMfmaOrders) {
            lds_a[sq][sk] = pa;
            lds_b[sq][sk] = pb;
        }
        double2 tc[8];
#pragma unroll
        for (int st = 0; st < 8; ++st) tc[st] = tn[st];
        __syncthreads();
        // The next item, and its loads.
        int64_t g1 = g0;
        int kc1 = kc + kMfmaOrders, nmax1 = nmax;
        if (kc1 >= nmax) {
            g1 = g0 + gstep;
            kc1 = 0;
            nmax1 = g1 < n_radius ? group_nmax(g1) : 0;
        }
        if (g1 < n_radius) load(g1, kc1, nmax1, pa, pb, tn);
        const int kend = nmax - kc < kMfmaOrders ? nmax - kc : kMfmaOrders;
#pragma unroll
        for (int st = 0; st < 8; ++st) {
            const int k4 = 4 * st;
            if (k4 < kend) {
                const int kk = k4 + kq;              // order index in the chunk
                const bool odd = (kc + kk + 1) & 1;  // the order's parity
                const double2 pt = tc[st];
                const double g1v = odd ? pt.x : pt.y;
                const double g2v = odd ? pt.y : pt.x;
                // The lane's A element, component comp of (u, v) = (a_n, b_n) at odd n, (b_n, a_n)
                // at even n, read from its own address (no per-lane branch).
                const double *src = reinterpret_cast<const double *>(
                                        (comp < 2) == odd ? &lds_a[0][0] : &lds_b[0][0]) +
                                    (comp & 1);
#pragma unroll
                for (int tt = 0; tt < 4; ++tt) {
                    const int q = 4 * tt + qin;
                    const double val = src[2 * (q * kMfmaOrders + kk)];
                    acc1[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(val, g1v, acc1[tt], 0, 0, 0);
                    acc2[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(val, g2v, acc2[tt], 0, 0, 0);
                }
            }
        }
        if (g1 == g0) {
            kc = kc1;
            continue;
        }
        // Epilogue of the group: node 4 tt + kq at cosine jj, all four components in the lane.
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            const int64_t node = g0 + 4 * tt + kq;
            const double cfac = node < n_radius ? tabc[gbase + node] : 0.0;
            const cx A = mk(acc1[tt][0], acc1[tt][1]), C = mk(acc1[tt][2], acc1[tt][3]);
            const cx D = mk(acc2[tt][0], acc2[tt][1]), B = mk(acc2[tt][2], acc2[tt][3]);
#pragma unroll
            for (int sgn = 0; sgn < 2; ++sgn) {
                const cx s1 = sgn == 0 ? A + B : A - B;
                const cx s2 = sgn == 0 ? C + D : C - D;
                const cx pp = s2 * conj(s1);
                tp[sgn][0] += cfac * (0.5 * (norm2(s1) + norm2(s2)));
                tp[sgn][1] += cfac * (0.5 * (norm2(s2) - norm2(s1)));
                tp[sgn][2] += cfac * pp.re;
                tp[sgn][3] += cfac * pp.im;
            }
        }
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            acc1[tt] = mfma_d4{0, 0, 0, 0};
            acc2[tt] = mfma_d4{0, 0, 0, 0};
        }
        g0 = g1;
        kc = kc1;
        nmax = nmax1;
    }
    // Add the four node lanes (kq = 0..3) of each cosine, in a fixed order.
#pragma unroll
    for (int sgn = 0; sgn < 2; ++sgn)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double x = tp[sgn][k];
            x += __shfl_xor(x, 16, 64);
            x += __shfl_xor(x, 32, 64);
            tp[sgn][k] = x;
        }
    if (jj < h) {
        // Lane kq writes component kq at +mu and -mu.
        double *out = tab_out + ((int64_t)sp * ne + eb) * 4 * (int64_t)n_angles;
        const int jp = n_angles - h + jj, jm = n_angles - 1 - jp;
        double vp = tp[0][0], vm = tp[1][0];
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            vp = kq == k ? tp[0][k] : vp;
            vm = kq == k ? tp[1][k] : vm;
        }
        out[kq * n_angles + jp] = vp;
        if (jm != jp) out[kq * n_angles + jm] = vm;
    }
}

// The angular products on the fp16 matrix cores (v_mfma_f32_16x16x32_f16) with both operands
// split in two fp16 slices, g = g_0 + g_1 and x = x_0 + x_1 (g_0, x_0 the nearest fp16 values,
// g_1, x_1 the fp16 values nearest the remainders), and G X = G_0 X_0 + G_0 X_1 + G_1 X_0 up to
// the dropped G_1 X_1, of order 2^-22 of the terms. The three products of a table run into one
// fp32 accumulator over kF16Run 32-order chunks, which is then added into fp64: its rounding,
// 2^-24 of the partial sums, sets the error floor (a few 1e-7 of the largest phase-function
// value over the 4096 x 64 lognormal workload). So that the remainders stay within fp16's normal
// range, each node's coefficients are scaled by 2^s (s its coefficient exponent: the largest
// near 2^10) and the table by 2^t (the largest entry, (2n + 1)/2 at most, near 2^15); the node's
// table factor carries 2^-2(s + t). All scalings are exact.
// The coefficients are the A operand and the table the B operand, C^T = X^T G^T: row 4 q + c of
// a group's X^T is node q's (u.re, u.im, v.re, v.im)[c] (four MFMA tiles of 4 nodes), a
// wavefront's 16 columns its cosines. v_mfma_f32_16x16x32_f16 leaves D[i][j] in lane
// j + 16 (i / 4), element i % 4, so each lane receives the four components of node 4 tt + lane/16
// at its cosine from both tables, A = G1 u, C = G1 v, D = G2 u, B = G2 v, and forms
// S1(+-mu) = A +- B, S2(+-mu) = C +- D and F11, F12, F33, F34 itself. The table slices are
// prepared once per call (pt_table_f16), the coefficients split while they are staged in LDS.
// A workgroup is one ensemble and 16 W non-negative cosines, groups of 16 nodes as
// ens_angular_mfma; each lane sums its nodes in group order and the four node lanes of a cosine
// are added at the end, always in the same order.
typedef _Float16 mfma_h8 __attribute__((ext_vector_type(8)));
typedef float mfma_f4 __attribute__((ext_vector_type(4)));

constexpr int kF16Orders = 32;   // one v_mfma_f32_16x16x32_f16 K step
constexpr int kF16Run = 4;       // chunks summed in fp32 before each fp64 flush
constexpr int kF16Pitch = 40;    // LDS row of 32 orders padded to 40 halves (bank spread)

// The exponent t of the table's scale 2^t: its entries are bounded by (2 n_max + 1)/2.
static int table_exponent(int n_max) {
    const double gmax = (2.0 * n_max + 1.0) / 2.0;
    return (int)std::floor(std::log2(32768.0 / gmax));
}

// tab16[(2 t + q) * h * npad + jj * npad + n - 1]: slice q of table t (G1, G2) times 2^texp;
// zero beyond n_max.
// The split partial tables of ens_angular_mfma added in split order: out[i] = sum_sp part[sp][i].
static __global__ void ens_split_reduce(const double *part, int splits, int64_t cells,
                                        double *out) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= cells) return;
    double acc = part[i];
    for (int sp = 1; sp < splits; ++sp) acc += part[(int64_t)sp * cells + i];
    out[i] = acc;
}

static __global__ void pt_table_f16(const double2 *table, int h, int n_max, int npad, int texp,
                                    _Float16 *tab16) {
    const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (idx >= (int64_t)h * npad) return;
    const int jj = (int)(idx / npad), k = (int)(idx % npad);  // k = n - 1
    double g[2] = {0.0, 0.0};
    if (k < n_max) {
        const double2 pt = table[(size_t)k * h + jj];
        const bool odd = (k + 1) & 1;
        g[0] = ldexp(odd ? pt.x : pt.y, texp);
        g[1] = ldexp(odd ? pt.y : pt.x, texp);
    }
    const size_t plane = (size_t)h * npad;
    const size_t o = (size_t)jj * npad + k;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
        const _Float16 g0 = (_Float16)g[t];
        tab16[(size_t)(2 * t) * plane + o] = g0;
        tab16[(size_t)(2 * t + 1) * plane + o] = (_Float16)(g[t] - (double)g0);
    }
}

// The loop runs over the items (group of 16 nodes, 32-order chunk) of the ensemble, ahead of
// the products: coefficients are loaded two items ahead and table slices one, issued in every
// step, table slices first (the memory counter retires in issue order, so the next step's wait
// for its table slices leaves these coefficients outstanding), from addresses clamped into
// valid data and masked after arrival. The ensemble's node list (N and coefficient exponent,
// staging offsets, table factors) and each group's largest N are copied to LDS at the start, so
// that moving to the next group issues no global loads. Dynamic LDS: f16_lds_bytes(n_radius);
// n_radius is at most kF16MaxNodes.
constexpr int64_t kF16MaxNodes = 2048;  // node list within 33 KB of LDS

__host__ __device__ inline size_t f16_lds_bytes(int64_t n_radius) {
    return (size_t)n_radius * (sizeof(double) + 2 * sizeof(int)) +
           (size_t)((n_radius + kGroupNodes - 1) / kGroupNodes) * sizeof(int);
}

// An item of the fp16 angular kernel: group g (its first node) and chunk kc (its first order).
struct Item {
    int g, kc;
};

// A workgroup is W wavefronts and 16 W cosines (a cosine slice); each thread stages P = 8 / W of
// an item's 512 (node, order) pairs.
constexpr int kF16Waves = 8;

template <int W>
static __global__ __launch_bounds__(64 * W) void ens_angular_f16(
    int64_t e0, int64_t ne, int64_t n_radius, int n_angles, int h, int slices, int npad,
    int texp, const int *node_n, const int64_t *node_off, const int64_t *ens_base,
    const float4 *staging, const double *tabc, const int *xexp, const _Float16 *tab16,
    double *tab_out) {
    constexpr int P = 8 / W;
    __shared__ _Float16 lds_x[2][2][64][kF16Pitch];  // [buffer][slice][row 4 q + c][order]
    extern __shared__ double w_cf[];  // [n_radius] table factors times 2^-2(s + t), then:
    int *w_n = reinterpret_cast<int *>(w_cf + n_radius);  // [n_radius] N + (s << 24)
    int *w_off = w_n + n_radius;                           // [n_radius] staging offsets
    int *gmax = w_off + n_radius;                          // [groups] largest N
    // Block b is slice b % slices of ensemble b / slices. The dispatcher deals blocks round-robin
    // over the 8 XCDs, so with a power-of-two number of slices each XCD serves fixed slices and
    // its L2 holds only their part of the table (the larger operand stream of this kernel).
    const int64_t eb = blockIdx.x / slices;
    const int slice = (int)(blockIdx.x % slices);
    const int64_t e = e0 + eb;
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int col = lane & 15, kq = lane >> 4;  // B column (cosine), order set; D node in tile
    // Staging pair p of this thread: node sq + p * 2 W of the group, order sk of the chunk.
    const int sq = t / kF16Orders, sk = t % kF16Orders;
    const int jj = slice * 16 * W + wave * 16 + col;  // this lane's cosine
    const bool col_ok = jj < h;
    const size_t plane = (size_t)h * npad;
    // Columns beyond h read row 0 of the table: finite values whose products are not written.
    const _Float16 *b_col = tab16 + (size_t)(col_ok ? jj : 0) * npad + 8 * kq;
    const float4 *stg = staging + ens_base[e];
    const int nr = (int)n_radius;
    const int ngroups = (nr + kGroupNodes - 1) / kGroupNodes;
    {
        const int *nn = node_n + e * n_radius;
        const int64_t *noff = node_off + e * n_radius;
        const double *tc = tabc + eb * n_radius;
        const int *xe = xexp + eb * n_radius;
        for (int i = t; i < ngroups; i += blockDim.x) gmax[i] = 0;
        for (int i = t; i < nr; i += blockDim.x) {
            const int sx = xe[i];
            w_cf[i] = ldexp(tc[i], -2 * (sx + texp));
            w_n[i] = nn[i] | (sx << 24);  // N < 2^24, s < 128
            w_off[i] = (int)noff[i];
        }
        __syncthreads();
        for (int i = t; i < nr; i += blockDim.x)
            atomicMax(&gmax[i / kGroupNodes], w_n[i] & 0xffffff);
        __syncthreads();
    }
    // A slot of staged coefficients: this thread's (node, order) pairs of one item (the fp32
    // (a.re, a.im, b.re, b.im)), loaded from addresses clamped into the ensemble's staged terms,
    // and the nodes' exponents; a pair beyond its node's N (or the ensemble's nodes) gets the
    // exponent -200, which takes its finite values to exactly zero in fp32.
    struct Slot {
        float4 v[P];
        int sx[P];
    };
    auto fetch = [&](Item it, Slot &sl) __attribute__((always_inline)) {
        const int g = it.g < nr ? it.g : 0;  // items past the end load valid, unused data
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const int node = g + sq + p * 2 * W;
            const bool ok = node < nr;
            const int packed = ok ? w_n[node] : 0;
            const int nq = packed & 0xffffff;
            const float4 *cp = stg + (ok ? w_off[node] : 0);
            const int n = it.kc + sk;  // n - 1
            const bool in = n < nq;
            sl.sx[p] = in ? packed >> 24 : -200;
            sl.v[p] = cp[in ? n : 0];
        }
    };
    // The slot's pairs of item `it` to LDS buffer `buf`: u, v by parity, times 2^s, in two slices
    // (in fp32: the scaling and the remainder x - x_0 are exact).
    auto convert = [&](Item it, const Slot &sl, int buf) __attribute__((always_inline)) {
        const bool odd = (it.kc + sk + 1) & 1;
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const float4 q = sl.v[p];
            const int e2 = sl.sx[p];
            const float xs[4] = {ldexpf(odd ? q.x : q.z, e2), ldexpf(odd ? q.y : q.w, e2),
                                 ldexpf(odd ? q.z : q.x, e2), ldexpf(odd ? q.w : q.y, e2)};
            const int row = 4 * (sq + p * 2 * W);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const _Float16 x0 = (_Float16)xs[c];
                lds_x[buf][0][row + c][sk] = x0;
                lds_x[buf][1][row + c][sk] = (_Float16)(xs[c] - (float)x0);
            }
        }
    };
    auto table = [&](int kc, mfma_h8 (&g)[2][2]) __attribute__((always_inline)) {
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int q = 0; q < 2; ++q)
                g[tb][q] = *reinterpret_cast<const mfma_h8 *>(b_col + (size_t)(2 * tb + q) * plane + kc);
    };
    // The item after `it` (groups past the end stay there).
    auto next = [&](Item it) __attribute__((always_inline)) {
        if (it.g >= nr) return it;
        if (it.kc + kF16Orders < gmax[it.g / kGroupNodes]) return Item{it.g, it.kc + kF16Orders};
        return Item{it.g + kGroupNodes, 0};
    };

    const mfma_f4 zero = {0.f, 0.f, 0.f, 0.f};
    double T[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};  // [sign][2 F11, 2 F12, F33, F34]
    double acc[4][2][4];  // [tile][table][component]: the group's sums so far (set by its first flush)
    mfma_f4 run[4][2];    // fp32 sums of up to kF16Run chunks
    bool fresh = true;    // no flush yet in this group
#pragma unroll
    for (int tt = 0; tt < 4; ++tt) run[tt][0] = run[tt][1] = zero;
    // Item cursors: c0 is computed, c1 converted (its table slices loaded), c2 waits in the
    // other slot, c3 is loaded into the slot that c1 frees.
    Item c0{0, 0};
    Item c1 = next(c0), c2 = next(c1), c3 = next(c2);
    Slot sa, sb;
    fetch(c0, sb);
    convert(c0, sb, 0);
    fetch(c1, sa);
    fetch(c2, sb);
    mfma_h8 gA[2][2], gB[2][2];
    table(c0.kc, gA);
    // One step: c1's table slices into gn, c0's products from LDS buffer bc with the table
    // slices gc alongside c1's split from slot `sl` into buffer bn, c3's loads into `sl` and, at
    // the end of a group, the epilogue. The loop alternates the slots, the table registers and
    // the LDS buffers, so that no copy waits for a load in flight and one barrier per step
    // suffices: buffer bn was last read one step ago, before every wavefront passed this step's
    // barrier, and buffer bc was written then.
    auto step = [&](Slot &sl, mfma_h8 (&gc)[2][2], mfma_h8 (&gn)[2][2], int bc, int bn)
                    __attribute__((always_inline)) {
        __syncthreads();
        table(c1.g < nr ? c1.kc : 0, gn);
        __builtin_amdgcn_sched_barrier(0);  // the table slices go out before anything else
        const bool same = c1.g == c0.g;
        // c0's products and c1's split into the other LDS buffer.
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            const mfma_h8 x0 = *reinterpret_cast<const mfma_h8 *>(&lds_x[bc][0][16 * tt + col][8 * kq]);
            const mfma_h8 x1 = *reinterpret_cast<const mfma_h8 *>(&lds_x[bc][1][16 * tt + col][8 * kq]);
#pragma unroll
            for (int tb = 0; tb < 2; ++tb) {
                mfma_f4 r = __builtin_amdgcn_mfma_f32_16x16x32_f16(x0, gc[tb][0], run[tt][tb], 0, 0, 0);
                r = __builtin_amdgcn_mfma_f32_16x16x32_f16(x0, gc[tb][1], r, 0, 0, 0);
                run[tt][tb] = __builtin_amdgcn_mfma_f32_16x16x32_f16(x1, gc[tb][0], r, 0, 0, 0);
            }
        }
        convert(c1, sl, bn);
        __builtin_amdgcn_sched_barrier(0);  // the slot is read before it is refilled
        fetch(c3, sl);
        // Into fp64 every kF16Run chunks and at the group's end; the group's first flush sets.
        if (!same || (c0.kc / kF16Orders) % kF16Run == kF16Run - 1) {
            if (fresh) {
#pragma unroll
                for (int tt = 0; tt < 4; ++tt)
#pragma unroll
                    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[tt][tb][v] = (double)run[tt][tb][v];
            } else {
#pragma unroll
                for (int tt = 0; tt < 4; ++tt)
#pragma unroll
                    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
                        for (int v = 0; v < 4; ++v) acc[tt][tb][v] += (double)run[tt][tb][v];
            }
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) run[tt][0] = run[tt][1] = zero;
            fresh = false;
        }
        if (!same) {
            // Epilogue of the group: node 4 tt + kq at cosine jj.
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) {
                const int node = c0.g + 4 * tt + kq;
                const double cfac = node < nr ? w_cf[node] : 0.0;
                const cx A = mk(acc[tt][0][0], acc[tt][0][1]), C = mk(acc[tt][0][2], acc[tt][0][3]);
                const cx D = mk(acc[tt][1][0], acc[tt][1][1]), B = mk(acc[tt][1][2], acc[tt][1][3]);
#pragma unroll
                for (int sgn = 0; sgn < 2; ++sgn) {
                    const cx s1 = sgn == 0 ? A + B : A - B;
                    const cx s2 = sgn == 0 ? C + D : C - D;
                    const double n1 = norm2(s1), n2 = norm2(s2);
                    const cx pp = s2 * conj(s1);
                    T[sgn][0] += cfac * (n1 + n2);
                    T[sgn][1] += cfac * (n2 - n1);
                    T[sgn][2] += cfac * pp.re;
                    T[sgn][3] += cfac * pp.im;
                }
            }
            fresh = true;
        }
        c0 = c1;
        c1 = c2;
        c2 = c3;
        c3 = next(c3);
        return c0.g < nr;
    };
    while (step(sa, gA, gB, 0, 1) && step(sb, gB, gA, 1, 0)) {
    }
    // Add the four node lanes (kq = 0..3) of the cosine, in a fixed order.
#pragma unroll
    for (int sgn = 0; sgn < 2; ++sgn)
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double x = T[sgn][k];
            x += __shfl_xor(x, 16, 64);
            x += __shfl_xor(x, 32, 64);
            T[sgn][k] = x;
        }
    if (col_ok) {
        // Lane kq writes component kq at +mu and -mu.
        double *out = tab_out + eb * 4 * (int64_t)n_angles;
        const int jp = n_angles - h + jj, jm = n_angles - 1 - jp;
        double vp = T[0][0], vm = T[1][0];
#pragma unroll
        for (int k = 1; k < 4; ++k) {
            vp = kq == k ? T[0][k] : vp;
            vm = kq == k ? T[1][k] : vm;
        }
        const double f = kq < 2 ? 0.5 : 1.0;
        out[kq * n_angles + jp] = f * vp;
        if (jm != jp) out[kq * n_angles + jm] = f * vm;
    }
}

// Averages and outputs of one ensemble (ensemble/engine.rs:1782-1836,2396-2497): scalars in
// SCALARS order, the phase matrix and, for the Greek projection (ens_greek_mfma), the factor
// gfac[eb] its coefficients take: 1/Z, or with normalization 1/alpha1_0 in the units of the
// unnormalized sums, alpha1_0 = (1/2) sum_j w_j F11(mu_j) d00_0(mu_j) (greek.rs:499-509;
// engine.rs:2447-2458), summed per thread in angle order and then over the threads in a fixed
// tree, so the result does not depend on scheduling.
static __global__ __launch_bounds__(256) void ens_finish(int64_t e0, int64_t n_radius, int n_angles,
                                                  const double *aw, const double *vals,
                                                  const double *tab, const double *dtab,
                                                  int64_t order, int normalized,
                                                  double *scalars, double *gfac,
                                                  double *phase, double *phase_fac) {
    __shared__ double part[kNodeVals][33];  // per value: 32 partial sums (padded row)
    __shared__ double S[kNodeVals];
    const int64_t e = e0 + blockIdx.x;
    const int t = threadIdx.x;
    const double *v = vals + blockIdx.x * n_radius * kNodeVals;
    {
        // The 8 node sums: thread t adds value t % 8 over the nodes i = t / 8 (mod 32), in node
        // order; the 32 partial sums of each value are then added in a fixed order.
        const int k = t % kNodeVals, r = t / kNodeVals;
        double a = 0.0;
        for (int64_t i = r; i < n_radius; i += 32) a += v[i * kNodeVals + k];
        part[k][r] = a;
        __syncthreads();
        if (t < kNodeVals) {
            double s = 0.0;
            for (int j = 0; j < 32; ++j) s += part[t][j];
            S[t] = s;
        }
        __syncthreads();
    }
    const double z = S[0], iz = 1.0 / z;
    const double *te = tab + (int64_t)blockIdx.x * 4 * n_angles;
    const double ext = S[1] * iz, sca = S[2] * iz, scag = S[3] * iz, absb = S[4] * iz,
                 back = S[5] * iz, r2 = S[6] * iz, r3 = S[7] * iz;
    if (t == 0) {
        double *o = scalars + e * 11;
        o[0] = ext;
        o[1] = sca;
        o[2] = absb;
        o[3] = back;
        o[4] = sca / ext;
        o[5] = scag / sca;
        o[6] = 4.0 * kPi * (ext / back);
        o[7] = 0.0;
        o[8] = r3 / r2;
        o[9] = kPi * r2;
        o[10] = 4.0 * kPi / 3.0 * r3;
    }
    // The phase matrix's factor: 1/Z, divided by <C_sca> when normalized; to phase_fac for the
    // phase at the caller's cosines (ens_phase_mu).
    if (phase_fac && t == 0) phase_fac[blockIdx.x] = normalized ? iz / sca : iz;
    if (phase) {
        // [F11, F12, F22 = F11, F33, F34, F44 = F33] / Z, divided by <C_sca> when normalized
        // (ensemble/engine.rs:2459-2490).
        const double f = normalized ? iz / sca : iz;
        double *ph = phase + e * 6 * (int64_t)n_angles;
        const int src[6] = {0, 1, 0, 2, 3, 2};
        for (int k = 0; k < 6; ++k)
            for (int j = t; j < n_angles; j += blockDim.x)
                ph[k * n_angles + j] = te[src[k] * n_angles + j] * f;
    }
    if (!gfac) return;
    const int64_t len = order + 1;
    __shared__ double red[256];
    double a = 0.0;
    for (int j = t; j < n_angles; j += blockDim.x) a += (aw[j] * te[j]) * dtab[(int64_t)j * len];
    red[t] = a;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (t < w) red[t] += red[t + w];
        __syncthreads();
    }
    if (t == 0) gfac[blockIdx.x] = normalized ? 1.0 / (0.5 * red[0]) : iz;
}

// The phase matrix at the caller's cosines (the reference's mu argument,
// python/scatrs/ensemble.py:550-562): tab holds each ensemble's summed tables at the
// symmetric set of n_out = 2 h cosines -|mu| and +|mu| (the distinct |mu| ascending), idx[j] the
// position of cosine j in it; phase[e][k][j] = tab[src_k][idx[j]] times the ensemble's factor
// from ens_finish, in [F11, F12, F22 = F11, F33, F34, F44 = F33] order.
static __global__ void ens_phase_mu(int64_t e0, int64_t ne, int n_out, const int *idx, int n_mu,
                                    const double *tab, const double *fac, double *phase) {
    const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (g >= ne * 6 * (int64_t)n_mu) return;
    const int64_t eb = g / (6 * (int64_t)n_mu);
    const int k = (int)((g / n_mu) % 6), j = (int)(g % n_mu);
    const int src = k == 0 || k == 2 ? 0 : k == 1 ? 1 : k == 4 ? 3 : 2;
    phase[(e0 + eb) * 6 * (int64_t)n_mu + (int64_t)k * n_mu + j] =
        tab[eb * 4 * (int64_t)n_out + (int64_t)src * n_out + idx[j]] * fac[eb];
}

// The Greek coefficients of a chunk's ensembles on the FP64 matrix cores (greek.rs:499-560;
// engine.rs:2396-2458). Per ensemble e and order l the six projections
// s_k(l) = sum_j A_k[e][j] D_k[j][l] of the weighted summed tables, A = w_j times
// (F11, F33, F11 + F33, F11 - F33, F12, F34), on the Wigner d functions
// D = (d00, d00, d22, d2-2, d02, d02) at the Gauss-Legendre cosines, as GEMMs in tiles of 16
// ensembles x 16 orders with v_mfma_f64_16x16x4f64 (A[i][k] from lane i + 16 k, B[k][j] from
// lane j + 16 k, D[i][j] in lane j + 16 (i % 4), element i / 4). A workgroup is 16 ensembles and
// 64 orders, a wavefront 16 orders. Then alpha1 = c s11, alpha2, alpha3 = c (sp +- sm)/2,
// alpha4 = c s33, beta1 = c s12, beta2 = c s34 with c = (2l + 1)/2 gfac[e]; alpha1_0 = 1 exactly
// when normalized. dtab[(p * n_angles + j) * len + l], p = d00, d22, d2-2, d02.
constexpr int kGreekAngles = 64;  // angles of the A tile staged per step

static __global__ __launch_bounds__(256) void ens_greek_mfma(int64_t e0, int64_t ne, int n_angles,
                                                            const double *aw, const double *tab,
                                                            const double *dtab, int64_t len,
                                                            const double *gfac, int normalized,
                                                            double *greek) {
    // The A tile: [combination][ensemble][angle], the six weighted combinations of 16
    // ensembles at kGreekAngles angles, staged with loads along each ensemble's rows.
    __shared__ double lds_a[6][16][kGreekAngles + 1];
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int r = lane & 15, kq = lane >> 4;
    const int64_t et = (int64_t)blockIdx.x * 16;               // the tile's first ensemble
    const int64_t lt = (int64_t)blockIdx.y * 64 + wave * 16;   // the wavefront's first order
    const int64_t lb = lt + r;                                 // this lane's B column
    const double *dcol = dtab + (lb < len ? lb : 0);
    const int64_t plane = (int64_t)n_angles * len;
    mfma_d4 acc[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) acc[k] = mfma_d4{0, 0, 0, 0};
    for (int j0 = 0; j0 < n_angles; j0 += kGreekAngles) {
        __syncthreads();
        // Thread t stages (ensemble t / 16, angles t % 16 + 16 h) of the tile.
#pragma unroll
        for (int h = 0; h < kGreekAngles / 16; ++h) {
            const int er = t / 16, ja = t % 16 + 16 * h;
            const int64_t e = et + er;
            const int j = j0 + ja;
            double f11 = 0.0, f12 = 0.0, f33 = 0.0, f34 = 0.0;
            if (e < ne && j < n_angles) {  // rows beyond the chunk, angles beyond the rule: 0
                const double *te = tab + e * 4 * (int64_t)n_angles + j;
                const double w = aw[j];
                f11 = w * te[0];
                f12 = w * te[n_angles];
                f33 = w * te[2 * n_angles];
                f34 = w * te[3 * n_angles];
            }
            lds_a[0][er][ja] = f11;
            lds_a[1][er][ja] = f33;
            lds_a[2][er][ja] = f11 + f33;
            lds_a[3][er][ja] = f11 - f33;
            lds_a[4][er][ja] = f12;
            lds_a[5][er][ja] = f34;
        }
        __syncthreads();
#pragma unroll 4
        for (int jj = 0; jj < kGreekAngles; jj += 4) {
            const int ja = jj + kq;
            const int j = j0 + ja;
            // Angles beyond the rule meet zero A values; read a valid row of the table for them.
            const double *dj = dcol + (int64_t)(j < n_angles ? j : 0) * len;
            const double d00 = dj[0], d22 = dj[plane], d2m2 = dj[2 * plane], d02 = dj[3 * plane];
            acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_a[0][r][ja], d00, acc[0], 0, 0, 0);
            acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_a[1][r][ja], d00, acc[1], 0, 0, 0);
            acc[2] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_a[2][r][ja], d22, acc[2], 0, 0, 0);
            acc[3] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_a[3][r][ja], d2m2, acc[3], 0, 0, 0);
            acc[4] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_a[4][r][ja], d02, acc[4], 0, 0, 0);
            acc[5] = __builtin_amdgcn_mfma_f64_16x16x4f64(lds_a[5][r][ja], d02, acc[5], 0, 0, 0);
        }
    }
    if (lb >= len) return;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        const int64_t e = et + kq + 4 * v;
        if (e >= ne) continue;
        const double c = (double)(2 * lb + 1) * 0.5 * gfac[e];
        const double sp = acc[2][v], sm = acc[3][v];
        double g6[6] = {c * acc[0][v], c * 0.5 * (sp + sm), c * 0.5 * (sp - sm), c * acc[1][v],
                        c * acc[4][v], c * acc[5][v]};
        if (normalized && lb == 0) g6[0] = 1.0;
        double *gr = greek + (e0 + e) * 6 * len + lb;
#pragma unroll
        for (int k = 0; k < 6; ++k) gr[k * len] = g6[k];
    }
}

// ---- host side ----

struct Rule {
    std::vector<double> nodes, weights;
};

// Uploads host data to a grow-only device buffer.
template <class T>
static int upload(scatrs_ctx *ctx, int which, const std::vector<T> &v, T **out) {
    int rc = check_hip(ctx, ensure(ctx, which, v.size() * sizeof(T)), "alloc");
    if (rc) return rc;
    *out = static_cast<T *>(ctx->buf[which]);
    if (ctx->replay.mode == 2) return SCATRS_OK;  // uploaded by the recording run
    return check_hip(ctx, hipMemcpyAsync(*out, v.data(), v.size() * sizeof(T),
                                         hipMemcpyHostToDevice, ctx->stream),
                     "upload");
}

// The reference's message for invalid distribution parameters (distribution.rs:99-185), or
// "" when they are valid.
static std::string dist_message(int kind, const double *p) {
    if (dist_valid(kind, p)) return "";
    const double p0 = p[0], p1 = p[1], p2 = p[2];
    const auto positive = [&](const char *name, double v) {
        return std::string(name) + " must be finite and > 0, got " + fmt_e(v);
    };
    const bool ok0 = std::isfinite(p0) && p0 > 0.0;
    if (kind == kLogNormal) {
        if (!ok0) return positive("median_radius", p0);
        return "the lognormal width must be finite and > 1, got " + fmt_e(p1);
    }
    if (kind == kGamma) {
        if (!ok0) return positive("effective_radius", p0);
        return "effective_variance must be in (0, 0.5), got " + fmt_e(p1);
    }
    if (kind == kModifiedGamma) {
        if (!(std::isfinite(p0) && p0 > -1.0))
            return "alpha must be finite and > -1, got " + fmt_e(p0);
        if (!(std::isfinite(p1) && p1 > 0.0)) return positive("b", p1);
        return positive("gamma", p2);
    }
    if (kind == kPowerLaw) {
        if (!std::isfinite(p0)) return "exponent must be finite, got " + fmt_e(p0);
        if (!(std::isfinite(p1) && p1 > 0.0)) return positive("min_radius", p1);
        if (!(std::isfinite(p2) && p2 > 0.0)) return positive("max_radius", p2);
        return "min_radius must be < max_radius, got " + fmt_e(p1) + " and " + fmt_e(p2);
    }
    if (kind == kMonodisperse) return positive("radius", p0);
    if (kind == kSegment) {
        if (!(std::isfinite(p0) && p0 > 0.0)) return positive("radius", p0);
        if (!(std::isfinite(p1) && p1 > p0)) return "tabulated radii must increase";
        return "a tabulated segment's tilt must be in [0, 1], got " + fmt_e(p2);
    }
    return "unknown distribution kind " + std::to_string(kind);
}

static std::string ensemble_message(scatrs_ctx *ctx, int64_t g, const EnsArgs &dev, int64_t n_radius,
                             const std::vector<double> &rnodes, int allow, int *code_out) {
    const int64_t e = g / n_radius, i = g % n_radius;
    const int L = dev.layers;
    int32_t kind = 0;
    double p[3] = {0.0, 0.0, 0.0}, wl = 0;
    std::vector<double> fr(std::max(L - 1, 0)), m(2 * L);
    if (!dev.user_rw) {
        (void)hipMemcpy(&kind, dev.kind + e, sizeof kind, hipMemcpyDeviceToHost);
        (void)hipMemcpy(&p[0], dev.p0 + e, sizeof(double), hipMemcpyDeviceToHost);
        (void)hipMemcpy(&p[1], dev.p1 + e, sizeof(double), hipMemcpyDeviceToHost);
        if (dev.p2) (void)hipMemcpy(&p[2], dev.p2 + e, sizeof(double), hipMemcpyDeviceToHost);
    }
    const double p0 = p[0], p1 = p[1], p2 = p[2];
    (void)hipMemcpy(&wl, dev.wl + e, sizeof wl, hipMemcpyDeviceToHost);
    if (L > 1)
        (void)hipMemcpy(fr.data(), dev.fr + e * (L - 1), fr.size() * sizeof(double),
                        hipMemcpyDeviceToHost);
    (void)hipMemcpy(m.data(), dev.m + e * 2 * L, m.size() * sizeof(double), hipMemcpyDeviceToHost);
    const std::string pre = "ensemble " + std::to_string(e) + ": ";
    *code_out = kInvalid;
    double2 rw = make_double2(0.0, 0.0);
    if (dev.user_rw) (void)hipMemcpy(&rw, dev.user_rw + g, sizeof rw, hipMemcpyDeviceToHost);
    if (!dev.user_rw && !dist_valid(kind, p)) return pre + dist_message(kind, p);
    if (!(std::isfinite(wl) && wl > 0.0))
        return pre + "the wavelength must be finite and > 0, got " + fmt_e(wl);
    for (int l = 0; l + 1 < L; ++l)
        if (!(fr[l] > 0.0 && fr[l] <= 1.0) || (l > 0 && fr[l] < fr[l - 1]))
            return pre + "the radius fractions must be nondecreasing in (0, 1]";
    if (dev.user_rw && !(std::isfinite(rw.x) && rw.x > 0.0 && std::isfinite(rw.y) && rw.y >= 0.0))
        return pre + "node " + std::to_string(i) +
               ": the radius must be finite and > 0 and the weight finite and >= 0, got " +
               fmt_e(rw.x) + " and " + fmt_e(rw.y);
    double t_lo = 0.0, t_hi = 0.0;
    if (!dev.user_rw) radius_range(kind, p, &t_lo, &t_hi);
    const double r = dev.user_rw ? rw.x
                                 : std::exp(0.5 * (t_hi + t_lo) + 0.5 * (t_hi - t_lo) * rnodes[i]);
    const double x = 2.0 * kPi * r / wl;
    const std::string at = pre + "at wavelength " + fmt_e(wl) + ", radius " + fmt_e(r) + ": ";
    std::string msg;
    if (L == 1) {
        *code_out = sphere_message(x, m[0], m[1], 0, allow, &msg);
        return at + msg;
    }
    *code_out = kOutside;
    return at + "a layered sphere of " + std::to_string(L) + " layers with x_L = " + fmt_e(x) +
           " is outside the validated domain (x_L in [1e-6, 1e4], |m| x_L <= 3e4, every "
           "layer's 0 < Re m <= 10, 0 <= Im m <= 10, |m| >= 0.1), or its inputs are invalid";
}

// Common checks of the two ensemble entry points' arguments.
static int ens_check_args(scatrs_ctx *ctx, int64_t count, int layers, int64_t n_radius,
                          int64_t n_angles, bool greek, int64_t order) {
    if (count < 0) return fail(ctx, SCATRS_INVALID_INPUT, "the number of ensembles must be >= 0");
    if (layers < 1 || layers > kMaxLayersDevice)
        return fail(ctx, SCATRS_INVALID_INPUT,
                    "layers must be in [1, " + std::to_string(kMaxLayersDevice) + "], got " +
                        std::to_string(layers));
    if (n_radius < 1 || n_radius > 65536)
        return fail(ctx, SCATRS_INVALID_INPUT, "n_radius must be in [1, 65536]");
    if (n_angles < 2 || n_angles > 4096)
        return fail(ctx, SCATRS_INVALID_INPUT, "n_angles must be in [2, 4096]");
    if (greek && (order < 0 || order > 65536))
        return fail(ctx, SCATRS_INVALID_INPUT, "the Greek order must be in [0, 65536]");
    return SCATRS_OK;
}

// What both entry points set up before their chunk loops: the Gauss-Legendre rules (host, then
// uploaded), per-node N and offsets, the validated status, the w pi_n, w tau_n table and, for
// Greek outputs, the Wigner d table.
struct EnsSetup {
    Rule rr, ar;
    int h = 0;
    double *d_rtab = nullptr, *d_atab = nullptr, *d_dtab = nullptr;
    double *d_dtab_t = nullptr;  // the Wigner d table transposed: [(p * len + l) * n_angles + j]
    EnsArgs a{};
    unsigned long long *counter = nullptr;
    int *node_n = nullptr;
    int64_t *node_off = nullptr, *ens_terms = nullptr, *ens_base = nullptr;
    std::vector<int64_t> terms, base;
    int n_max = 0;
    int64_t total_terms = 0;
    double2 *table = nullptr;
    _Float16 *tab16 = nullptr;  // the fp16-split table (angular kernel 2)
    int npad = 0;               // its rows' length: n_max rounded up to 32
    int texp = 0;               // its scale 2^texp
    int kernel = 0;             // the angular kernel of the call
    bool stage32 = false;       // the node solves stage the fp32 form (kernel 2, forward calls)
    bool values_only = false;   // a forward call without angular products: no staging needed
    // A forward call (scatrs_ensemble_fixed / _nodes): its node plan serves only its own node
    // launches, so small launches may solve one node per wavefront. The reverse pass reuses its
    // forward's plan for the reverse node kernels, which need the slab layout.
    bool forward_only = false;
};

// The angle rule of n cosines from the context's cache, computed and uploaded on a miss (the
// least recently used of kMaxAngleRules entries makes room).
static int angle_rule(scatrs_ctx *ctx, int64_t n, scatrs_ctx::AngleRule **out) {
    int rc;
    ++ctx->use_clock;
    for (auto &r : ctx->angle_rules)
        if (r.n == n) {
            r.used = ctx->use_clock;
            *out = &r;
            return SCATRS_OK;
        }
    if ((int)ctx->angle_rules.size() >= scatrs_ctx::kMaxAngleRules) {
        auto old = std::min_element(ctx->angle_rules.begin(), ctx->angle_rules.end(),
                                    [](const auto &x, const auto &y) { return x.used < y.used; });
        if ((rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync"))) return rc;
        if (old->d_atab) (void)hipFree(old->d_atab);
        if (old->d_dtab) (void)hipFree(old->d_dtab);
        ctx->angle_rules.erase(old);
    }
    scatrs_ctx::AngleRule r;
    r.n = n;
    r.used = ctx->use_clock;
    gauss_legendre(n, r.nodes, r.weights);
    const int h = (int)((n + 1) / 2);
    std::vector<double> atab(n + h);  // weights, then the non-negative cosines
    std::copy(r.weights.begin(), r.weights.end(), atab.begin());
    for (int jj = 0; jj < h; ++jj) atab[n + jj] = r.nodes[n - h + jj];
    if ((rc = check_hip(ctx, hipMalloc(&r.d_atab, atab.size() * sizeof(double)), "alloc")))
        return rc;
    if ((rc = check_hip(ctx, hipMemcpyAsync(r.d_atab, atab.data(), atab.size() * sizeof(double),
                                            hipMemcpyHostToDevice, ctx->stream),
                        "upload"))) {
        (void)hipFree(r.d_atab);
        return rc;
    }
    ctx->angle_rules.push_back(std::move(r));
    *out = &ctx->angle_rules.back();
    return SCATRS_OK;
}

static int ens_setup(scatrs_ctx *ctx, int64_t count, const int32_t *d_kind, const double *d_p0,
                     const double *d_p1, const double *d_p2, const double *d_wavelength,
                     int layers,
                     const double *d_fractions, const double *d_m, int64_t n_radius,
                     int64_t n_angles, bool greek, int64_t order, int allow_unvalidated,
                     int kernel, EnsSetup &S, const double *d_nodes = nullptr,
                     bool rule_tables = true) {
    int rc;
    if ((rc = check_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice"))) return rc;
    hipStream_t s = ctx->stream;
    const int h = (int)((n_angles + 1) / 2);
    S.h = h;
    // The Gauss-Legendre rule in radius (device copy in kMisc) is kept in the context for the
    // last size, the angle rules for the last kMaxAngleRules sizes (the host rules cost
    // O(n^2)). The caller's nodes need no radius rule (a one-node stand-in).
    const int64_t nr_rule = d_nodes ? 1 : n_radius;
    if (ctx->rule_nr != nr_rule) {
        ctx->rule_nr = -1;
        gauss_legendre(nr_rule, ctx->rr_nodes, ctx->rr_weights);
        std::vector<double> rtab(2 * nr_rule);
        std::copy(ctx->rr_nodes.begin(), ctx->rr_nodes.end(), rtab.begin());
        std::copy(ctx->rr_weights.begin(), ctx->rr_weights.end(), rtab.begin() + nr_rule);
        double *dr = nullptr;
        if ((rc = upload(ctx, scatrs_ctx::kMisc, rtab, &dr))) return rc;
        ctx->rule_nr = nr_rule;
    }
    scatrs_ctx::AngleRule *A = nullptr;
    if ((rc = angle_rule(ctx, n_angles, &A))) return rc;
    S.rr.nodes = ctx->rr_nodes;
    S.rr.weights = ctx->rr_weights;
    S.ar.nodes = A->nodes;
    S.ar.weights = A->weights;
    S.d_rtab = static_cast<double *>(ctx->buf[scatrs_ctx::kMisc]);
    S.d_atab = A->d_atab;
    S.a = EnsArgs{d_kind, d_p0, d_p1, d_wavelength, d_fractions, d_m, layers, n_radius,
                  S.d_rtab, S.d_rtab + nr_rule, d_p2};
    // Per-node N and status.
    const int64_t total = count * n_radius;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kCounters, sizeof(Status) + 64), "alloc")))
        return rc;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kKeys, (size_t)total * sizeof(int)), "alloc")))
        return rc;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kPerm,
                                    (size_t)total * sizeof(int64_t) + (size_t)count * 2 * sizeof(int64_t)),
                        "alloc")))
        return rc;
    Status *st = static_cast<Status *>(ctx->buf[scatrs_ctx::kCounters]);
    S.counter = reinterpret_cast<unsigned long long *>(
        static_cast<char *>(ctx->buf[scatrs_ctx::kCounters]) + sizeof(Status));
    S.node_n = static_cast<int *>(ctx->buf[scatrs_ctx::kKeys]);
    S.node_off = static_cast<int64_t *>(ctx->buf[scatrs_ctx::kPerm]);
    S.ens_terms = S.node_off + total;
    S.ens_base = S.ens_terms + count;
    const int64_t pblocks = (total + 255) / 256;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kHist, (size_t)pblocks * sizeof(Partial)),
                        "alloc")))
        return rc;
    Partial *partial = static_cast<Partial *>(ctx->buf[scatrs_ctx::kHist]);
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kGeom, (size_t)total * sizeof(double2)),
                        "alloc")))
        return rc;
    S.a.geom = static_cast<double2 *>(ctx->buf[scatrs_ctx::kGeom]);
    S.a.user_rw = reinterpret_cast<const double2 *>(d_nodes);
    ens_prep<<<(unsigned)pblocks, 256, 0, s>>>(S.a, count, allow_unvalidated, S.node_n, partial);
    auto &rp = ctx->replay;
    if (rp.mode == 2) {
        // A replay computes with the recorded N at every node (the launches were planned for
        // them) and flags any node whose N its inputs changed.
        fill_u32<<<1, 1, 0, s>>>(reinterpret_cast<unsigned *>(rp.guard), 0u, 1);
        plan_guard<<<(unsigned)pblocks, 256, 0, s>>>(S.node_n, rp.node_n_rec, total, rp.guard);
    }
    reduce_status<<<1, 1024, 0, s>>>(partial, pblocks, st);
    ens_offsets<<<(unsigned)count, 256, 0, s>>>(S.node_n, n_radius, S.node_off, S.ens_terms);
    if (rp.mode == 2) {
        S.terms = rp.terms;
        S.base.assign(count, 0);
        S.total_terms = rp.total_terms;
        S.n_max = rp.n_max;
    }
    Status *hst = static_cast<Status *>(ctx->host_pinned);
    // The status and each ensemble's number of terms come back to the host (not in a replay).
    if (rp.mode != 2) {
        if ((rc = check_hip(ctx, hipMemcpyAsync(hst, st, sizeof(Status), hipMemcpyDeviceToHost, s),
                            "status copy")))
            return rc;
        S.terms.assign(count, 0);
        S.base.assign(count, 0);
        if ((rc = check_hip(ctx, hipMemcpyAsync(S.terms.data(), S.ens_terms, count * sizeof(int64_t),
                                                hipMemcpyDeviceToHost, s),
                            "terms copy")))
            return rc;
        if ((rc = check_hip(ctx, hipStreamSynchronize(s), "ensemble prep"))) return rc;
        const unsigned long long bad = std::min(hst->first_bad, hst->first_outside);
        if (bad != ~0ull) {
            int code = kInvalid;
            const std::string msg = ensemble_message(ctx, (int64_t)bad, S.a, n_radius, S.rr.nodes,
                                                     allow_unvalidated, &code);
            return fail(ctx, code == kOutside ? SCATRS_OUTSIDE_VALIDATED_DOMAIN : SCATRS_INVALID_INPUT,
                        msg);
        }
        S.total_terms = (int64_t)hst->total_terms;
        S.n_max = hst->n_max;
        if (rp.mode == 1) {
            rp.terms = S.terms;
            rp.total_terms = S.total_terms;
            rp.n_max = S.n_max;
            if ((rc = check_hip(ctx, hipMalloc(&rp.node_n_rec, (size_t)total * sizeof(int) + 256),
                                "alloc")))
                return rc;
            rp.guard = reinterpret_cast<int *>(reinterpret_cast<char *>(rp.node_n_rec) +
                                               (size_t)total * sizeof(int));
            if ((rc = check_hip(ctx, hipMemcpyAsync(rp.node_n_rec, S.node_n, (size_t)total * sizeof(int),
                                                    hipMemcpyDeviceToDevice, s),
                                "record")))
                return rc;
        }
    }
    // The w pi_n, w tau_n table at the non-negative cosines up to the largest N (not made when
    // the call takes no angular products at the rule's cosines: rule_tables false).
    if (rule_tables) {
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kTables,
                                        (size_t)S.n_max * h * sizeof(double2)),
                            "table")))
            return rc;
        S.table = static_cast<double2 *>(ctx->buf[scatrs_ctx::kTables]);
        pt_table<<<(h + 63) / 64, 64, 0, s>>>(S.d_atab + n_angles, h, S.n_max, S.table);
    }
    S.kernel = kernel == 2 && n_radius > kF16MaxNodes ? 0 : kernel;
    S.stage32 = S.kernel == 2;
    if (S.kernel == 2) {
        S.npad = (S.n_max + kF16Orders - 1) / kF16Orders * kF16Orders;
        S.texp = table_exponent(S.n_max);
    }
    if (S.kernel == 2 && rule_tables) {
        const size_t cells = (size_t)h * S.npad;
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kTables16, 4 * cells * sizeof(_Float16)),
                            "table")))
            return rc;
        S.tab16 = static_cast<_Float16 *>(ctx->buf[scatrs_ctx::kTables16]);
        pt_table_f16<<<(unsigned)((cells + 255) / 256), 256, 0, s>>>(S.table, h, S.n_max, S.npad,
                                                                     S.texp, S.tab16);
    }
    if (greek) {
        if (A->order != order) {
            A->order = -1;
            const int64_t len = order + 1;
            // Both layouts: angle-major for the Greek projection, order-major (after it) for
            // its transpose in reverse mode.
            const size_t cells = 4 * (size_t)n_angles * len;
            std::vector<double> dt(2 * cells);
            const int pm[4][2] = {{0, 0}, {2, 2}, {2, -2}, {0, 2}};
            for (int p = 0; p < 4; ++p)
                for (int64_t j = 0; j < n_angles; ++j) {
                    double *row = dt.data() + ((size_t)p * n_angles + j) * len;
                    wigner_d_small(pm[p][0], pm[p][1], S.ar.nodes[j], order, row);
                    for (int64_t l = 0; l < len; ++l)
                        dt[cells + ((size_t)p * len + l) * n_angles + j] = row[l];
                }
            const size_t bytes = dt.size() * sizeof(double);
            if (A->dtab_cap < bytes) {
                if (A->d_dtab) {
                    if ((rc = check_hip(ctx, hipStreamSynchronize(s), "sync"))) return rc;
                    (void)hipFree(A->d_dtab);
                    A->d_dtab = nullptr;
                    A->dtab_cap = 0;
                }
                if ((rc = check_hip(ctx, hipMalloc(&A->d_dtab, bytes), "alloc"))) return rc;
                A->dtab_cap = bytes;
            }
            if ((rc = check_hip(ctx, hipMemcpyAsync(A->d_dtab, dt.data(), bytes,
                                                    hipMemcpyHostToDevice, s),
                                "upload")))
                return rc;
            A->order = order;
        }
        S.d_dtab = A->d_dtab;
        S.d_dtab_t = S.d_dtab + 4 * n_angles * (order + 1);
    }
    return SCATRS_OK;
}

// The next chunk [e0, e1) of ensembles whose coefficients fit max_terms staged terms (and at
// most 2^26 nodes); fills S.base for them (staging offsets relative to e0) and returns e1.
static int64_t next_chunk(EnsSetup &S, int64_t e0, int64_t count, int64_t n_radius,
                          int64_t max_terms, int64_t *chunk_terms) {
    int64_t e1 = e0, acc = 0;
    while (e1 < count && (e1 == e0 || acc + S.terms[e1] <= max_terms) &&
           (e1 - e0) * n_radius < (int64_t)1 << 26) {
        S.base[e1] = acc;
        acc += S.terms[e1];
        ++e1;
    }
    *chunk_terms = acc;
    return e1;
}

// A launch of the node kernels: the ranked nodes [start, end), slabs of `rows` rows; `waves`
// is the persistent launch's number of wavefronts (spheres), 0 for a loop-free launch.
struct NodeLaunch {
    int64_t start, end, rows, waves;
};

// Buffers of one forward chunk.
struct ChunkBufs {
    cx *staging = nullptr;
    const float4 *staging32 = nullptr;  // the fp32 form the fp16-split kernel reads
    double *vals = nullptr, *tabc = nullptr, *tab_out = nullptr;
    int *xexp = nullptr;             // per node: its coefficient exponent
    int *order = nullptr;            // the chunk's nodes ranked by N, largest first
    std::vector<NodeLaunch> plan;    // the node-kernel launches over that ranking
};

// Scratch budgets of the node launches' slabs: a loop-free layered launch (and its largest
// number of wavefronts), and the persistent sphere launch, whose slabs all take the chunk's
// largest N.
constexpr size_t kSlabBudget = (size_t)4 << 30;
constexpr int64_t kMaxNodeWaves = 16384;
constexpr size_t kSphereSlabBudget = (size_t)16 << 30;

// Staging buffer of a chunk's coefficients (term_bytes per term: 32 in FP64, 16 in fp32); it
// bounds the chunk size. Every chunk waits for its slowest node (the serial recurrence of the
// largest N), so chunks are made as large as memory allows: 1/`share` of the device's free
// memory, at most 64 GiB and at least 1 GiB. Returns the staged terms a chunk may hold.
static int64_t chunk_terms_limit(int share, size_t term_bytes) {
    size_t free_b = 0, total_b = 0;
    size_t bytes = (size_t)12 << 30;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && free_b > 0)
        bytes = std::min<size_t>((size_t)64 << 30, std::max<size_t>((size_t)1 << 30, free_b / share));
    return (int64_t)(bytes / term_bytes);
}

template <int L>
static void launch_nodes_ranked(hipStream_t s, const EnsArgs &a, int64_t e0, const EnsSetup &S,
                                cx *scratch, const ChunkBufs &B, cx *back_out) {
    for (const NodeLaunch &q : B.plan) {
        // (A one-node-per-wavefront plan, q.waves < 0, serves tangent mode too: ens_nodes_wave.)
        if (a.tan_dir >= 0 && q.waves >= 0) {
            if constexpr (L == 1) {
                fill_u64<<<1, 1, 0, s>>>(S.counter, 0ull, 1);
                ens_nodes_tangent_persist<<<(unsigned)(q.waves / 4), 256, 0, s>>>(
                    a, e0, q.end, B.order, S.node_n, S.node_off, S.ens_base, S.counter, scratch,
                    q.rows, B.staging, S.stage32, B.vals, B.tabc, B.xexp);
            } else {
                ens_nodes_tangent<L><<<(unsigned)((q.end - q.start + 255) / 256), 256, 0, s>>>(
                    a, e0, B.order, q.start, q.end, S.node_n, S.node_off, S.ens_base, scratch,
                    q.rows, B.staging, S.stage32, B.vals, B.tabc, B.xexp);
            }
            continue;
        }
        if constexpr (L == 1) {
            fill_u64<<<1, 1, 0, s>>>(S.counter, 0ull, 1);
            if (q.waves < 0) {
                ens_nodes_wave<<<(unsigned)(-q.waves), 64, 0, s>>>(
                    a, e0, q.end, B.order, S.node_n, S.node_off, S.ens_base, S.counter,
                    reinterpret_cast<char *>(scratch), q.rows, mie_wave_bytes(q.rows - 2),
                    S.values_only ? nullptr : B.staging, S.stage32, B.vals, B.tabc, B.xexp,
                    back_out);
                continue;
            }
            ens_nodes_persist<<<(unsigned)(q.waves / 4), 256, 0, s>>>(
                a, e0, q.end, B.order, S.node_n, S.node_off, S.ens_base, S.counter, scratch,
                q.rows, B.staging, S.stage32, B.vals, B.tabc, B.xexp, back_out);
        } else {
            ens_nodes_ranked<L><<<(unsigned)((q.end - q.start + 255) / 256), 256, 0, s>>>(
                a, e0, B.order, q.start, q.end, S.node_n, S.node_off, S.ens_base, scratch,
                q.rows, B.staging, S.stage32, B.vals, B.tabc, B.xexp, back_out);
        }
    }
}

// Keeps a chunk's node-launch plan for the replay (recording run only).
static void record_plan(scatrs_ctx *ctx, const ChunkBufs &B, size_t scratch_bytes) {
    auto &rp = ctx->replay;
    if (rp.mode != 1) return;
    std::vector<int64_t> q;
    for (const NodeLaunch &l : B.plan) q.insert(q.end(), {l.start, l.end, l.rows, l.waves});
    rp.plans.push_back(std::move(q));
    rp.scratch.push_back(scratch_bytes);
}

// Ranks the chunk's nodes by N, largest first, into B.order, and plans the node launches of
// B.plan over that ranking. Spheres: one persistent launch, slabs sized by the chunk's largest
// N, as many wavefronts as kSphereSlabBudget holds (at most 8 per CU). Layered spheres:
// consecutive ranges, each with slabs sized by its first (largest) node's N and as many
// wavefronts as kSlabBudget holds (at most kMaxNodeWaves), so that a few nodes of large N do not
// shrink the launches of the many small ones. *scratch_bytes receives the slab bytes of the
// largest launch. Synchronizes: the histogram of N comes back to the host.
// The node launches of a chunk (nodes ranked by N on the device). Spheres: one persistent
// launch whose slabs are sized by the chunk's largest N, from the N histogram or, for a chunk
// that is the whole call (`whole`), the call's largest N without reading the histogram back.
static int plan_nodes(scatrs_ctx *ctx, const EnsSetup &S, int64_t e0, int64_t nodes,
                      int64_t chunk_terms, bool whole, ChunkBufs &B, size_t *scratch_bytes,
                      hipStream_t s) {
    int rc;
    const int kord = scatrs_ctx::kOrder;
    if ((rc = check_hip(ctx, ensure(ctx, kord,
                                    (size_t)nodes * 2 * sizeof(int) + 2 * kBuckets * sizeof(unsigned)),
                        "order")))
        return rc;
    int *keys = static_cast<int *>(ctx->buf[kord]);
    B.order = keys + nodes;
    unsigned *hist = reinterpret_cast<unsigned *>(B.order + nodes);
    fill_u32<<<(2 * kBuckets + 255) / 256, 256, 0, s>>>(hist, 0u, 2 * kBuckets);
    ens_node_keys<<<(unsigned)((nodes + 255) / 256), 256, 0, s>>>(e0, S.a.n_radius, nodes, S.node_n,
                                                                 keys);
    histogram_keys<<<(unsigned)((nodes + kSortBlock - 1) / kSortBlock), 256, 0, s>>>(keys, nodes,
                                                                                    hist);
    sort_by_keys(keys, hist, B.order, nodes, s);
    auto &rp = ctx->replay;
    if (rp.mode == 2) {
        if (rp.next_plan >= rp.plans.size())
            return fail(ctx, SCATRS_DEVICE_ERROR, "the graph capture met more chunks than recorded");
        const std::vector<int64_t> &q = rp.plans[rp.next_plan];
        B.plan.clear();
        for (size_t i = 0; i + 3 < q.size(); i += 4)
            B.plan.push_back(NodeLaunch{q[i], q[i + 1], q[i + 2], q[i + 3]});
        *scratch_bytes = rp.scratch[rp.next_plan++];
        return SCATRS_OK;
    }
    std::vector<unsigned> h(kBuckets);
    const bool sphere_whole = S.a.layers == 1 && whole;
    if (!sphere_whole) {
        if ((rc = check_hip(ctx, hipMemcpyAsync(h.data(), hist, kBuckets * sizeof(unsigned),
                                                hipMemcpyDeviceToHost, s),
                            "histogram copy")))
            return rc;
        if ((rc = check_hip(ctx, hipStreamSynchronize(s), "node ranking"))) return rc;
    }
    B.plan.clear();
    *scratch_bytes = 0;
    if (S.a.layers == 1) {
        int b0 = 0;
        while (!sphere_whole && b0 < kBuckets - 1 && h[b0] == 0) ++b0;
        // Bucket 0 holds every N above kBuckets - 2: its rows come from the largest N (the
        // first nonempty bucket's N is the chunk's largest, the call's for a whole call).
        const int64_t n = (sphere_whole || b0 == 0) ? S.n_max : std::max(kBuckets - 2 - b0, 0);
        const int64_t rows = n + 2;
        if (S.forward_only &&
            wave_rule(ctx->tune[SCATRS_TUNE_SPHERE_WAVE], ctx->cus, chunk_terms, n)) {
            // One node per wavefront (ens_nodes_wave), two wavefronts per SIMD at most: the
            // plan's negative count marks it.
            const int64_t waves = std::min<int64_t>(nodes, (int64_t)ctx->cus * 8);
            B.plan.push_back(NodeLaunch{0, nodes, rows, -waves});
            *scratch_bytes = (size_t)(mie_wave_bytes(n) * waves);
            record_plan(ctx, B, *scratch_bytes);
            return SCATRS_OK;
        }
        const size_t per_wave = (size_t)slab_elems(1, rows) * sizeof(cx);
        int64_t waves = std::min<int64_t>((int64_t)ctx->cus * ctx->tune[SCATRS_TUNE_NODE_WAVES],
                                          (nodes + 63) / 64);
        waves = std::min<int64_t>(waves, std::max<int64_t>(4, (int64_t)(kSphereSlabBudget / per_wave)));
        waves = (waves + 3) / 4 * 4;
        B.plan.push_back(NodeLaunch{0, nodes, rows, waves});
        *scratch_bytes = per_wave * (size_t)waves;
        record_plan(ctx, B, *scratch_bytes);
        return SCATRS_OK;
    }
    int64_t pos = 0, below = 0;  // below: ranked nodes in the buckets before b
    int b = 0;
    while (pos < nodes) {
        while (b < kBuckets - 1 && below + h[b] <= pos) below += h[b++];
        // Bucket 0 holds every N above kBuckets - 2: its rows come from the largest N.
        const int64_t n = b == 0 ? S.n_max : std::max(kBuckets - 2 - b, 0);
        const int64_t rows = n + 2;
        const size_t per_wave = (size_t)slab_elems(S.a.layers, rows) * sizeof(cx);
        const int64_t waves = std::min<int64_t>(
            kMaxNodeWaves, std::max<int64_t>(4, (int64_t)(kSlabBudget / per_wave)));
        const int64_t end = std::min<int64_t>(nodes, pos + waves * 64);
        B.plan.push_back(NodeLaunch{pos, end, rows, 0});
        *scratch_bytes = std::max(*scratch_bytes, per_wave * (size_t)((end - pos + 255) / 256 * 4));
        pos = end;
    }
    record_plan(ctx, B, *scratch_bytes);
    return SCATRS_OK;
}

// The node solves of the chunk [e0, e0 + ne) (chunk_terms staged terms) on stream s:
// afterwards B.vals holds the nodes' weighted scalars, B.tabc their table
// factors, B.xexp their coefficient exponents and B.staging the coefficients. back_out (may be
// null) receives the nodes' backscatter sums. Synchronizes s (the node ranking).
static int forward_nodes(scatrs_ctx *ctx, EnsSetup &S, int64_t e0, int64_t ne,
                         int64_t chunk_terms, int64_t n_radius, int64_t n_angles, ChunkBufs &B,
                         cx *back_out, hipStream_t s) {
    int rc;
    const int layers = S.a.layers;
    const int64_t nodes = ne * n_radius;
    const int kstg = scatrs_ctx::kStaging, knod = scatrs_ctx::kNodes, kacc = scatrs_ctx::kAccum;
    const size_t term_bytes = S.stage32 ? sizeof(float4) : 2 * sizeof(cx);
    size_t scratch_bytes = 0;
    const bool whole = e0 == 0 && ne == (int64_t)S.terms.size();
    if ((rc = plan_nodes(ctx, S, e0, nodes, chunk_terms, whole, B, &scratch_bytes, s))) return rc;
    if ((rc = check_hip(ctx, ensure(ctx, kstg, (size_t)std::max<int64_t>(chunk_terms, 1) * term_bytes),
                        "staging")))
        return rc;
    if ((rc = check_hip(ctx, ensure(ctx, knod,
                                    (size_t)nodes * ((kNodeVals + 1) * sizeof(double) + sizeof(int))),
                        "nodes")))
        return rc;
    if ((rc = check_hip(ctx, ensure(ctx, kacc, (size_t)ne * 4 * n_angles * sizeof(double)),
                        "tables")))
        return rc;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kScratch, scratch_bytes), "scratch")))
        return rc;
    // ens_base holds the staging offsets of this chunk's ensembles, relative to e0's (a replay
    // finds them where the recording run put them).
    if (ctx->replay.mode != 2 &&
        (rc = check_hip(ctx, hipMemcpyAsync(S.ens_base + e0, S.base.data() + e0, ne * sizeof(int64_t),
                                            hipMemcpyHostToDevice, s),
                        "base copy")))
        return rc;
    B.staging = static_cast<cx *>(ctx->buf[kstg]);
    B.staging32 = S.stage32 ? reinterpret_cast<const float4 *>(B.staging) : nullptr;
    B.vals = static_cast<double *>(ctx->buf[knod]);
    B.tabc = B.vals + nodes * kNodeVals;
    B.xexp = reinterpret_cast<int *>(B.tabc + nodes);
    B.tab_out = static_cast<double *>(ctx->buf[kacc]);
    cx *scratch = static_cast<cx *>(ctx->buf[scatrs_ctx::kScratch]);
    mark_begin(ctx, s);
    const EnsArgs &a = S.a;
#define SCATRS_RANKED(LL) launch_nodes_ranked<LL>(s, a, e0, S, scratch, B, back_out)
    switch (layers) {
        case 1: SCATRS_RANKED(1); break;
        case 2: SCATRS_RANKED(2); break;
        case 3: SCATRS_RANKED(3); break;
        case 4: SCATRS_RANKED(4); break;
        case 5: SCATRS_RANKED(5); break;
        case 6: SCATRS_RANKED(6); break;
        case 7: SCATRS_RANKED(7); break;
        default: SCATRS_RANKED(8); break;
    }
#undef SCATRS_RANKED
    mark_end(ctx, s);
    return check_hip(ctx, hipGetLastError(), "ensemble launch");
}

// The fp32 form (a.re, a.im, b.re, b.im per term) of FP64-staged coefficients, for the
// fp16-split angular kernel where the FP64 coefficients are kept as well (reverse mode).
static __global__ void stage_fp32(const cx *staging, int64_t terms, float4 *out) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < terms;
         i += (int64_t)gridDim.x * blockDim.x) {
        const cx a = staging[2 * i], b = staging[2 * i + 1];
        out[i] = make_float4((float)a.re, (float)a.im, (float)b.re, (float)b.im);
    }
}

// The angular products of the chunk [e0, e0 + ne) on stream s from the buffers forward_nodes
// filled (B.staging32 with kernel 2): afterwards B.tab_out holds the ensembles' summed tables [F11, F12, F33, F34] x
// n_angles.
// The angular products with the call's kernel at h non-negative cosines (table, and its fp16
// slices for kernel 2), into tab_out laid out for n_out cosines (n_out - h + jj the cosine jj,
// n_out - 1 - that its mirror).
static int angular_products(scatrs_ctx *ctx, EnsSetup &S, int64_t e0, int64_t ne,
                            int64_t n_radius, int h, int n_out, const double2 *table,
                            const _Float16 *tab16, ChunkBufs &B, double *tab_out, hipStream_t s) {
    mark_begin(ctx, s);
    if (S.kernel == 1) {
        const int bt = (int)std::min<int64_t>(512, ((h + 63) / 64) * 64);
        ens_angular<<<(unsigned)ne, bt, 0, s>>>(e0, n_radius, n_out, h, S.node_n, S.node_off,
                                                S.ens_base, B.staging, B.tabc, table, tab_out);
    } else if (S.kernel == 2) {
        constexpr int W = kF16Waves;
        const int slices = (h + 16 * W - 1) / (16 * W);
        ens_angular_f16<W><<<(unsigned)(ne * slices), 64 * W, f16_lds_bytes(n_radius), s>>>(
            e0, ne, n_radius, n_out, h, slices, S.npad, S.texp, S.node_n, S.node_off, S.ens_base,
            B.staging32, B.tabc, B.xexp, tab16, tab_out);
    } else {
        const int slices = (h + kSliceAngles - 1) / kSliceAngles;
        // Small batches: node groups over more blocks, up to about two per CU in all.
        const int64_t groups = (n_radius + kGroupNodes - 1) / kGroupNodes;
        const int64_t want =
            (ctx->tune[SCATRS_TUNE_SPLIT_BLOCKS] * (int64_t)ctx->cus + ne * slices - 1) /
            (ne * slices);
        const int splits = (int)std::max<int64_t>(1, std::min<int64_t>({groups, want, 16}));
        const int64_t cells = ne * 4 * (int64_t)n_out;
        double *dst = tab_out;
        if (splits > 1) {
            int rc;
            if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kMisc3,
                                            (size_t)splits * cells * sizeof(double)),
                                "split tables")))
                return rc;
            dst = static_cast<double *>(ctx->buf[scatrs_ctx::kMisc3]);
        }
        const int lds_nodes = n_radius <= kMfmaLdsNodes;
        ens_angular_mfma<<<(unsigned)(ne * slices * splits), 512,
                           lds_nodes ? mfma_lds_bytes(n_radius) : 0, s>>>(
            e0, n_radius, n_out, h, slices, splits, ne, S.node_n, S.node_off, S.ens_base,
            B.staging, B.tabc, table, dst, lds_nodes);
        if (splits > 1)
            ens_split_reduce<<<(unsigned)((cells + 255) / 256), 256, 0, s>>>(dst, splits, cells,
                                                                           tab_out);
    }
    mark_end(ctx, s);
    return check_hip(ctx, hipGetLastError(), "ensemble launch");
}

static int forward_angular(scatrs_ctx *ctx, EnsSetup &S, int64_t e0, int64_t ne,
                           int64_t n_radius, int64_t n_angles, ChunkBufs &B, hipStream_t s) {
    return angular_products(ctx, S, e0, ne, n_radius, S.h, (int)n_angles, S.table, S.tab16, B,
                            B.tab_out, s);
}

// The caller's phase cosines of a call (scatrs_set_phase_cosines), on the device: the w pi_n,
// w tau_n table at the h distinct |mu| (and its fp16 slices with kernel 2), each cosine's
// place among the 2 h cosines -|mu|, +|mu|, and the cosines gathered by place.
struct PhaseMu {
    int h = 0, n_mu = 0;
    double2 *table = nullptr;
    _Float16 *tab16 = nullptr;
    int *idx = nullptr, *q_ptr = nullptr, *q_list = nullptr;
};

static int phase_mu_setup(scatrs_ctx *ctx, const EnsSetup &S, PhaseMu &U) {
    int rc;
    hipStream_t s = ctx->stream;
    U.h = (int)ctx->phase_amu.size();
    U.n_mu = (int)ctx->phase_mu.size();
    double *d_amu = nullptr;
    if ((rc = upload(ctx, scatrs_ctx::kPhaseMu, ctx->phase_amu, &d_amu))) return rc;
    const size_t tab_bytes = (size_t)S.n_max * U.h * sizeof(double2);
    const size_t cells = (size_t)U.h * S.npad;
    const size_t int_off = tab_bytes + (S.kernel == 2 ? 4 * cells * sizeof(_Float16) : 0);
    const size_t n_int = (size_t)U.n_mu + (2 * (size_t)U.h + 1) + (size_t)U.n_mu;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kTablesMu, int_off + n_int * sizeof(int)),
                        "table")))
        return rc;
    char *base = static_cast<char *>(ctx->buf[scatrs_ctx::kTablesMu]);
    U.table = reinterpret_cast<double2 *>(base);
    U.idx = reinterpret_cast<int *>(base + int_off);
    U.q_ptr = U.idx + U.n_mu;
    U.q_list = U.q_ptr + 2 * U.h + 1;
    const struct {
        int *dst;
        const std::vector<int> &src;
    } copies[3] = {{U.idx, ctx->phase_idx}, {U.q_ptr, ctx->phase_qptr}, {U.q_list, ctx->phase_qlist}};
    for (const auto &c : copies)
        if (ctx->replay.mode != 2 &&
            (rc = check_hip(ctx, hipMemcpyAsync(c.dst, c.src.data(), c.src.size() * sizeof(int),
                                                hipMemcpyHostToDevice, s),
                            "upload")))
            return rc;
    pt_table<<<(U.h + 63) / 64, 64, 0, s>>>(d_amu, U.h, S.n_max, U.table);
    if (S.kernel == 2) {
        U.tab16 = reinterpret_cast<_Float16 *>(base + tab_bytes);
        pt_table_f16<<<(unsigned)((cells + 255) / 256), 256, 0, s>>>(U.table, U.h, S.n_max,
                                                                     S.npad, S.texp, U.tab16);
    }
    return check_hip(ctx, hipGetLastError(), "table");
}

// The node solves and the angular products of a chunk on the context's stream.
static int forward_chunk(scatrs_ctx *ctx, EnsSetup &S, int64_t e0, int64_t ne,
                         int64_t chunk_terms, int64_t n_radius, int64_t n_angles, ChunkBufs &B,
                         cx *back_out) {
    int rc;
    if ((rc = forward_nodes(ctx, S, e0, ne, chunk_terms, n_radius, n_angles, B, back_out,
                            ctx->stream)))
        return rc;
    return forward_angular(ctx, S, e0, ne, n_radius, n_angles, B, ctx->stream);
}

}  // namespace scatrs
