// Single-particle batches beyond the sphere efficiencies: the sizing query, the coefficients
// a_n, b_n (and their derivatives) of spheres and layered spheres, and the efficiencies of
// layered spheres with every layer's Jacobian. One particle per lane; the particles are ranked
// by N (largest first) so that a wavefront's lanes run about the same number of orders, and
// solved by launches of one group of 64 per wavefront, each launch's scratch slabs sized by its
// first (largest) particle.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/scatrs_hip.h"
#include "checks.h"
#include "context.h"
#include "layered_device.h"
#include "mie_device.h"
#include "sort.h"
#include "util.h"

namespace scatrs {

// The sphere batch's wavefront-per-sphere solve of a checked, ranked batch (scatrs_hip.hip).
int mie_wave_for(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t len,
                 int64_t n_fixed, const int *perm, int64_t terms, int64_t nmax, double *d_values,
                 double *d_jacobian, bool *taken);

namespace {

constexpr int kModeEfficiencies = 0;
constexpr int kModeCoefficients = 1;

// Where a batch writes its results. Coefficients: a, b (count, width) complex, zero-padded
// beyond each particle's N; d (L, 4, count, width) [da/dx_l, da/dm_l, db/dx_l, db/dm_l] or null;
// terms (count) each particle's N. Efficiencies: values (5, count), jac (L, 3, 5, count) or null.
struct SingleOut {
    cx *a, *b, *d;
    int64_t *terms;
    int64_t width;
    double *values, *jac;
};

// Per particle: the checks, N, and its sort key; block partials for the status.
__global__ void __launch_bounds__(256) single_prep(const double *x, const double *m, int64_t count,
                                                   int L, int64_t n_fixed, int allow, int derivs,
                                                   int *keys, int *nterms, int64_t *terms_out,
                                                   Partial *partial) {
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    int n = 0;
    bool ok = false;
    unsigned long long bad = ~0ull, outside = ~0ull;
    if (i < count) {
        double xl[kMaxLayersDevice], ml[2 * kMaxLayersDevice];
        for (int l = 0; l < L; ++l) {
            xl[l] = x[i * L + l];
            ml[2 * l] = m[2 * (i * L + l)];
            ml[2 * l + 1] = m[2 * (i * L + l) + 1];
        }
        long long nn = 0;
        const int code = layered_x_code(xl, ml, L, n_fixed, allow, derivs, &nn);
        if (code == kInvalid) bad = (unsigned long long)i;
        if (code == kOutside) outside = (unsigned long long)i;
        ok = code == kValid;
        n = ok ? (int)nn : 0;
        keys[i] = bucket_of(n);
        nterms[i] = n;
        if (terms_out) terms_out[i] = n;
    }
    block_partial(n, ok, bad, outside, partial);
}

// Writes a particle's a_n, b_n and, with derivatives, every layer's da/dx_l, da/dm_l, db/dx_l,
// db/dm_l: the outer layer's from the order loop, the inner layers' from their stored
// sensitivities (layered/kernel.rs:549-587).
template <int L, bool DERIVS>
struct CoefSink {
    cx *a, *b;       // the particle's rows
    cx *d;           // its derivative row of (layer 0, kind 0)
    int64_t d_kind;  // elements between kinds (count x width)
    cx cur_a, cur_b;

    __device__ void order(int k, cx av, cx bv) {
        a[k - 1] = av;
        b[k - 1] = bv;
    }

    __device__ void put(int l, int k, cx dax, cx dam, cx dbx, cx dbm) {
        cx *p = d + (int64_t)l * 4 * d_kind + (k - 1);
        p[0] = dax;
        p[d_kind] = dam;
        p[2 * d_kind] = dbx;
        p[3 * d_kind] = dbm;
    }

    __device__ void order_d(int k, cx av, cx bv, cx dax, cx dam, cx dbx, cx dbm) {
        order(k, av, bv);
        put(L - 1, k, dax, dam, dbx, dbm);
    }

    template <class O>
    __device__ void stored(int k, cx inv_da, cx inv_db, cx, cx, cx, cx, const O &o) {
        if constexpr (L > 1) {
            const cx mL = o.m[L - 1];
            const cx inv_mL = inv(mL);
            const cx ida2 = inv_da * inv_da, idb2 = inv_db * inv_db;
#pragma unroll
            for (int l = 0; l < L - 1; ++l) {
                const cx ba_x = o.dh[l][0] * inv_mL, ba_m = o.dh[l][1] * inv_mL;
                const cx bb_x = mL * o.dh[l][2], bb_m = mL * o.dh[l][3];
                put(l, k, -(times_i(ba_x) * ida2), -(times_i(ba_m) * ida2),
                    -(times_i(bb_x) * idb2), -(times_i(bb_m) * idb2));
            }
        } else {
            (void)k, (void)inv_da, (void)inv_db, (void)o;
        }
    }

    __device__ void done(double, double, double, double) {}
    __device__ void zero_absorption() {}
};

// The efficiencies of a layered sphere and, with derivatives, the Jacobian of each layer's
// (x_l, Re m_l, Im m_l): the outer layer's coefficient derivatives come from the order loop,
// the inner layers' from their stored sensitivities, and each layer's absorption Jacobian as
// the ensemble adjoint forms it (layered/kernel.rs:549-587, mie/efficiencies.rs:89-146).
template <int L, bool DERIVS>
struct LayeredEffSink {
    EfficiencySink<false> v;
    DerivSums ds[DERIVS ? L : 1];
    cx cur_a, cur_b;
    double cur_own, cur_pair, cur_rk1;

    __device__ void order(int k, cx a, cx b) { v.order(k, a, b); }

    __device__ void order_d(int k, cx a, cx b, cx dax, cx dam, cx dbx, cx dbm) {
        const cx dak[3] = {dax, dam, times_i(dam)};
        const cx dbk[3] = {dbx, dbm, times_i(dbm)};
        v.weights(k, &cur_own, &cur_pair, &cur_rk1);
        ds[L - 1].add(k, (double)(2 * k + 1), cur_own, cur_pair, a, b, v.a_prev, v.b_prev, dak,
                      dbk);
        cur_a = a;
        cur_b = b;
        if constexpr (L == 1) v.order(k, a, b);
    }

    template <class O>
    __device__ void stored(int k, cx inv_da, cx inv_db, cx g_a, cx g_b, cx d_m, cx md,
                           const O &o) {
        if constexpr (L > 1 && DERIVS) {
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
                const cx dak[3] = {dax, dam, times_i(dam)};
                const cx dbk[3] = {dbx, dbm, times_i(dbm)};
                ds[l].add(k, w, cur_own, cur_pair, cur_a, cur_b, v.a_prev, v.b_prev, dak, dbk);
                ds[l].jac_abs[0] += w * (absorption_term(d_m, ba_x, ba_x * g_a, na) +
                                         absorption_term(md, bb_x, bb_x * g_b, nb));
                ds[l].jac_abs[1] += w * (absorption_term(d_m, ba_m, ba_m * g_a, na) +
                                         absorption_term(md, bb_m, bb_m * g_b, nb));
                const cx iba = times_i(ba_m), ibb = times_i(bb_m);
                ds[l].jac_abs[2] += w * (absorption_term(d_m, iba, iba * g_a, na) +
                                         absorption_term(md, ibb, ibb * g_b, nb));
            }
            v.order(k, cur_a, cur_b);
        } else {
            (void)k, (void)inv_da, (void)inv_db, (void)g_a, (void)g_b, (void)d_m, (void)md,
                (void)o;
        }
    }

    __device__ void done(double abs_s, double j0, double j1, double j2) {
        v.abs_sum = abs_s;
        if constexpr (DERIVS) {
            ds[L - 1].jac_abs[0] = j0;
            ds[L - 1].jac_abs[1] = j1;
            ds[L - 1].jac_abs[2] = j2;
        }
    }

    // A lossless layered sphere absorbs nothing, and stays lossless as any x_l or Re m_l
    // changes (layered/kernel.rs:453-464).
    __device__ void zero_absorption() {
        v.abs_sum = 0.0;
        if constexpr (DERIVS) {
#pragma unroll
            for (int l = 0; l < L; ++l) ds[l].jac_abs[0] = ds[l].jac_abs[1] = 0.0;
        }
    }
};

// Slab elements of one wavefront: spheres keep the checkpoint column and u_k(x) (as the
// ensemble node solves do); layered spheres keep 2 L - 1 full complex ratio columns and u_k(x).
__host__ __device__ inline int64_t single_slab_elems(int layers, int64_t rows) {
    if (layers == 1) return 64 * ((rows - 2) / kSphereBlock + 2 + (rows + 1) / 2);
    return rows * 64 * 2 * layers;
}

// One launch: ranked particles [start, end), one group of 64 per wavefront.
template <int L, bool DERIVS, int MODE>
__global__ __launch_bounds__(256) void single_solve(const double *__restrict__ x,
                                                    const double *__restrict__ m, int64_t count,
                                                    const int *__restrict__ perm,
                                                    const int *__restrict__ nterms, int64_t start,
                                                    int64_t end, cx *scratch, int64_t rows,
                                                    SingleOut out) {
    __shared__ cx ztile[4][L == 1 ? kSphereBlock * 64 : 1];
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    cx *slab = scratch + wave * single_slab_elems(L, rows) + lane;
    const int64_t idx = start + wave * 64 + lane;
    if (idx >= end) return;
    const int64_t i = perm[idx];
    const int n = nterms[i];
    const double *xi = x + i * L;
    const double *mi = m + 2 * i * L;
    auto solve = [&](auto &sink) {
        if constexpr (L == 1) {
            const cx mm = mk(mi[0], mi[1]);
            cx *tile = &ztile[threadIdx.x >> 6][lane];
            double *ux = real_column(slab, ((rows - 2) / kSphereBlock + 2) * 64);
            const auto zs = ratio_chain_ckpt<kSphereBlock>(mm, xi[0], n, slab, 64, tile);
            x_chain(xi[0], n, upward_orders(xi[0], n), ux, 64);
            NoOuter none;
            mie_orders<DERIVS>(xi[0], mm, n, zs, ux, 64, none, sink);
        } else {
            layered_solve_xl<L, DERIVS>(xi, mi, n, slab, rows, sink);
        }
    };
    if constexpr (MODE == kModeCoefficients) {
        CoefSink<L, DERIVS> sink;
        sink.a = out.a + i * out.width;
        sink.b = out.b + i * out.width;
        sink.d_kind = count * out.width;
        sink.d = DERIVS ? out.d + i * out.width : nullptr;
        solve(sink);
    } else if constexpr (L == 1) {
        // the sphere batch's own solve, so that one layer gives the sphere bit for bit
        double q[5], j[15];
        sphere_efficiencies<kSphereBlock, DERIVS>(
            xi[0], mk(mi[0], mi[1]), n, slab, real_column(slab, ((rows - 2) / kSphereBlock + 2) * 64),
            &ztile[threadIdx.x >> 6][lane], true, q, j);
#pragma unroll
        for (int k = 0; k < 5; ++k) out.values[k * count + i] = q[k];
        if constexpr (DERIVS) {
#pragma unroll
            for (int k = 0; k < 15; ++k) out.jac[k * count + i] = j[k];
        }
    } else {
        LayeredEffSink<L, DERIVS> sink;
        solve(sink);
        double q[5];
        sink.v.values(xi[L - 1], q);
#pragma unroll
        for (int k = 0; k < 5; ++k) out.values[k * count + i] = q[k];
        if constexpr (DERIVS) {
#pragma unroll
            for (int l = 0; l < L; ++l) {
                double j[15];
                sink.ds[l].jacobian(xi[L - 1], q, sink.v.sca, sink.v.back, l == L - 1, j);
#pragma unroll
                for (int k = 0; k < 15; ++k) out.jac[((int64_t)l * 15 + k) * count + i] = j[k];
            }
        }
    }
}

template <int L, bool DERIVS, int MODE>
void launch_one(int64_t blocks, hipStream_t s, const double *x, const double *m, int64_t count,
                const int *perm, const int *nterms, int64_t start, int64_t end, cx *scratch,
                int64_t rows, const SingleOut &out) {
    single_solve<L, DERIVS, MODE><<<(unsigned)blocks, 256, 0, s>>>(x, m, count, perm, nterms,
                                                                   start, end, scratch, rows,
                                                                   out);
}

template <bool DERIVS, int MODE>
bool launch_layers(int L, int64_t blocks, hipStream_t s, const double *x, const double *m,
                   int64_t count, const int *perm, const int *nterms, int64_t start, int64_t end,
                   cx *scratch, int64_t rows, const SingleOut &out) {
    switch (L) {
#define SCATRS_SINGLE_L(NL)                                                                    \
    case NL:                                                                                   \
        launch_one<NL, DERIVS, MODE>(blocks, s, x, m, count, perm, nterms, start, end, scratch,  \
                                     rows, out);                                               \
        return true;
        SCATRS_SINGLE_L(1)
        SCATRS_SINGLE_L(2)
        SCATRS_SINGLE_L(3)
        SCATRS_SINGLE_L(4)
        SCATRS_SINGLE_L(5)
        SCATRS_SINGLE_L(6)
        SCATRS_SINGLE_L(7)
        SCATRS_SINGLE_L(8)
#undef SCATRS_SINGLE_L
        default:
            return false;
    }
}

// The whole call: checks and sizing (one synchronization), then, unless sizing_only, the
// ranked solve in launches whose slabs fit the scratch budget.
int single_run(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t count, int L,
               int64_t n_max, int allow, int derivs, int mode, bool sizing_only,
               const SingleOut &out, int64_t *max_terms, int64_t *total_terms) {
    if (!ctx) return SCATRS_INVALID_INPUT;
    ctx->error.clear();
    if (max_terms) *max_terms = 0;
    if (total_terms) *total_terms = 0;
    if (count < 0) return fail(ctx, SCATRS_INVALID_INPUT, "the number of particles must be >= 0");
    if (L < 1 || L > kMaxLayersDevice)
        return fail(ctx, SCATRS_INVALID_INPUT,
                    "layers must be in [1, " + std::to_string(kMaxLayersDevice) + "], got " +
                        std::to_string(L));
    if (n_max < 0 || n_max > kMaxOrder)
        return fail(ctx, SCATRS_INVALID_INPUT,
                    "the number of terms must be in [1, 1048576], got " + std::to_string(n_max));
    if (count == 0) return SCATRS_OK;
    if (count > (int64_t)INT32_MAX)
        return fail(ctx, SCATRS_INVALID_INPUT, "at most 2^31 - 1 particles per call");
    if (!d_x || !d_m) return fail(ctx, SCATRS_INVALID_INPUT, "null input pointer");
    int rc;
    if ((rc = check_hip(ctx, hipSetDevice(ctx->device), "hipSetDevice"))) return rc;
    const int64_t blocks = (count + 255) / 256;
    if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kCounters, sizeof(Status) + 64), "alloc")) ||
        (rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kHist, 2 * kBuckets * sizeof(unsigned)),
                        "alloc")) ||
        (rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kKeys, (size_t)count * 2 * sizeof(int)),
                        "alloc")) ||
        (rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kPerm, (size_t)count * sizeof(int)), "alloc")) ||
        (rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kNodes, (size_t)blocks * sizeof(Partial)),
                        "alloc")))
        return rc;
    Status *st = static_cast<Status *>(ctx->buf[scatrs_ctx::kCounters]);
    unsigned *hist = static_cast<unsigned *>(ctx->buf[scatrs_ctx::kHist]);
    int *keys = static_cast<int *>(ctx->buf[scatrs_ctx::kKeys]);
    int *nterms = keys + count;
    int *perm = static_cast<int *>(ctx->buf[scatrs_ctx::kPerm]);
    Partial *partial = static_cast<Partial *>(ctx->buf[scatrs_ctx::kNodes]);
    hipStream_t s = ctx->stream;
    fill_u32<<<(2 * kBuckets + 255) / 256, 256, 0, s>>>(hist, 0u, 2 * kBuckets);
    single_prep<<<(unsigned)blocks, 256, 0, s>>>(d_x, d_m, count, L, n_max, allow, derivs, keys,
                                                  nterms, sizing_only ? nullptr : out.terms,
                                                  partial);
    reduce_status<<<1, 1024, 0, s>>>(partial, blocks, st);
    histogram_keys<<<(unsigned)((count + kSortBlock - 1) / kSortBlock), 256, 0, s>>>(keys, count,
                                                                                     hist);
    Status *hst = static_cast<Status *>(ctx->host_pinned);
    std::vector<unsigned> hh(kBuckets);
    if ((rc = check_hip(ctx, hipMemcpyAsync(hst, st, sizeof(Status), hipMemcpyDeviceToHost, s),
                        "status copy")) ||
        (rc = check_hip(ctx, hipMemcpyAsync(hh.data(), hist, kBuckets * sizeof(unsigned),
                                            hipMemcpyDeviceToHost, s),
                        "histogram copy")) ||
        (rc = check_hip(ctx, hipStreamSynchronize(s), "prep")))
        return rc;
    const unsigned long long bad = std::min(hst->first_bad, hst->first_outside);
    if (bad != ~0ull) {
        std::vector<double> hx(L), hm(2 * L);
        (void)hipMemcpy(hx.data(), d_x + bad * L, L * sizeof(double), hipMemcpyDeviceToHost);
        (void)hipMemcpy(hm.data(), d_m + 2 * bad * L, 2 * L * sizeof(double),
                        hipMemcpyDeviceToHost);
        std::string msg;
        const int code = layered_x_message(hx.data(), hm.data(), L, n_max, allow, derivs, &msg);
        return fail(ctx, code == kOutside ? SCATRS_OUTSIDE_VALIDATED_DOMAIN : SCATRS_INVALID_INPUT,
                    "element " + std::to_string(bad) + ": " + msg);
    }
    const int64_t nmax = hst->n_max;
    if (max_terms) *max_terms = nmax;
    if (total_terms) *total_terms = (int64_t)hst->total_terms;
    if (sizing_only) return SCATRS_OK;
    if (mode == kModeCoefficients) {
        if (!out.a || !out.b || !out.terms || (derivs && !out.d))
            return fail(ctx, SCATRS_INVALID_INPUT, "null output pointer");
        if (out.width < nmax)
            return fail(ctx, SCATRS_INVALID_INPUT,
                        "the output width " + std::to_string(out.width) +
                            " is below the largest number of terms " + std::to_string(nmax));
        const int64_t per = count * out.width * 2;  // doubles per array
        fill_f64<<<(unsigned)((per + 255) / 256), 256, 0, s>>>(reinterpret_cast<double *>(out.a),
                                                               0.0, per);
        fill_f64<<<(unsigned)((per + 255) / 256), 256, 0, s>>>(reinterpret_cast<double *>(out.b),
                                                               0.0, per);
        if (derivs) {
            const int64_t dn = per * 4 * L;
            fill_f64<<<(unsigned)((dn + 255) / 256), 256, 0, s>>>(
                reinterpret_cast<double *>(out.d), 0.0, dn);
        }
    } else if (!out.values || (derivs && !out.jac)) {
        return fail(ctx, SCATRS_INVALID_INPUT, "null output pointer");
    }
    sort_by_keys(keys, hist, perm, count, s);
    if (L == 1 && mode != kModeCoefficients) {
        // One layer takes the sphere batch's kernel choice, so that it gives the sphere bit for
        // bit: a small batch one sphere per wavefront.
        bool taken = false;
        if ((rc = mie_wave_for(ctx, d_x, d_m, count, n_max, perm, (int64_t)hst->total_terms, nmax,
                               out.values, derivs ? out.jac : nullptr, &taken)))
            return rc;
        if (taken) return SCATRS_OK;
    }
    // Launches over the ranked particles: each takes as many groups of 64 as slabs sized by its
    // first particle's N fit in the budget (a quarter of the free memory, at most 8 GiB).
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    const size_t budget = std::max<size_t>(size_t(256) << 20,
                                           std::min<size_t>(free_b / 4, size_t(8) << 30));
    int64_t start = 0;
    int bucket = 0;
    int64_t in_bucket = 0;  // ranked particles of `bucket` already covered
    while (start < count) {
        while (bucket < kBuckets && in_bucket >= (int64_t)hh[bucket]) {
            ++bucket;
            in_bucket = 0;
        }
        const int64_t n_first = bucket == 0 ? nmax : (int64_t)(kBuckets - 2 - bucket);
        const int64_t rows = std::max<int64_t>(n_first, 1) + 2;
        const size_t per_wave = (size_t)single_slab_elems(L, rows) * sizeof(cx);
        const int64_t waves = std::max<int64_t>(1, (int64_t)(budget / per_wave));
        const int64_t take = std::min<int64_t>(count - start, waves * 64);
        const int64_t lw = (take + 63) / 64;
        if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kScratch, per_wave * (size_t)lw),
                            "scratch")))
            return rc;
        cx *scratch = static_cast<cx *>(ctx->buf[scatrs_ctx::kScratch]);
        const int64_t nb = (lw + 3) / 4;
        bool ok;
        if (mode == kModeCoefficients)
            ok = derivs ? launch_layers<true, kModeCoefficients>(L, nb, s, d_x, d_m, count, perm,
                                                                 nterms, start, start + take,
                                                                 scratch, rows, out)
                        : launch_layers<false, kModeCoefficients>(L, nb, s, d_x, d_m, count, perm,
                                                                  nterms, start, start + take,
                                                                  scratch, rows, out);
        else
            ok = derivs ? launch_layers<true, kModeEfficiencies>(L, nb, s, d_x, d_m, count, perm,
                                                                 nterms, start, start + take,
                                                                 scratch, rows, out)
                        : launch_layers<false, kModeEfficiencies>(L, nb, s, d_x, d_m, count,
                                                                  perm, nterms, start,
                                                                  start + take, scratch, rows,
                                                                  out);
        if (!ok) return fail(ctx, SCATRS_INVALID_INPUT, "unsupported layer count");
        if ((rc = check_hip(ctx, hipGetLastError(), "single-particle launch"))) return rc;
        // advance the bucket cursor past the particles this launch took
        int64_t left = take;
        while (left > 0 && bucket < kBuckets) {
            const int64_t room = (int64_t)hh[bucket] - in_bucket;
            const int64_t used = std::min(room, left);
            in_bucket += used;
            left -= used;
            if (in_bucket >= (int64_t)hh[bucket]) {
                ++bucket;
                in_bucket = 0;
            }
        }
        start += take;
    }
    return SCATRS_OK;
}

}  // namespace
}  // namespace scatrs

using namespace scatrs;

extern "C" {

int scatrs_mie_terms(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t len,
                     int64_t n_max, int allow_unvalidated, int64_t *max_terms,
                     int64_t *total_terms) {
    SingleOut out{};
    return single_run(ctx, d_x, d_m, len, 1, n_max, allow_unvalidated, 0, kModeCoefficients,
                      true, out, max_terms, total_terms);
}

int scatrs_mie_coefficients(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t len,
                            int64_t n_max, int allow_unvalidated, int64_t width,
                            int64_t *d_terms, double *d_a, double *d_b, double *d_derivatives,
                            int64_t *total_terms) {
    SingleOut out{reinterpret_cast<cx *>(d_a), reinterpret_cast<cx *>(d_b),
                  reinterpret_cast<cx *>(d_derivatives), d_terms, width, nullptr, nullptr};
    return single_run(ctx, d_x, d_m, len, 1, n_max, allow_unvalidated, d_derivatives != nullptr,
                      kModeCoefficients, false, out, nullptr, total_terms);
}

int scatrs_layered_terms(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t count,
                         int layers, int64_t n_max, int allow_unvalidated, int derivatives,
                         int64_t *max_terms, int64_t *total_terms) {
    SingleOut out{};
    return single_run(ctx, d_x, d_m, count, layers, n_max, allow_unvalidated, derivatives != 0,
                      kModeCoefficients, true, out, max_terms, total_terms);
}

int scatrs_layered_coefficients(scatrs_ctx *ctx, const double *d_x, const double *d_m,
                                int64_t count, int layers, int64_t n_max, int allow_unvalidated,
                                int64_t width, int64_t *d_terms, double *d_a, double *d_b,
                                double *d_derivatives, int64_t *total_terms) {
    SingleOut out{reinterpret_cast<cx *>(d_a), reinterpret_cast<cx *>(d_b),
                  reinterpret_cast<cx *>(d_derivatives), d_terms, width, nullptr, nullptr};
    return single_run(ctx, d_x, d_m, count, layers, n_max, allow_unvalidated,
                      d_derivatives != nullptr, kModeCoefficients, false, out, nullptr,
                      total_terms);
}

int scatrs_layered_efficiencies(scatrs_ctx *ctx, const double *d_x, const double *d_m,
                                int64_t count, int layers, int64_t n_max, int allow_unvalidated,
                                double *d_values, double *d_jacobian, int64_t *total_terms) {
    SingleOut out{nullptr, nullptr, nullptr, nullptr, 0, d_values, d_jacobian};
    return single_run(ctx, d_x, d_m, count, layers, n_max, allow_unvalidated,
                      d_jacobian != nullptr, kModeEfficiencies, false, out, nullptr,
                      total_terms);
}

}  // extern "C"
