// The ensemble path on the fixed discrete grid (scatrs_ensemble_fixed): per ensemble (one
// distribution at one wavelength), n_radius Gauss-Legendre nodes in ln r, a Mie solve per node,
// the phase matrix at the n_angles Gauss-Legendre cosines by the parity-split angular
// products, the quadrature-weighted reduction, and the Greek projection. The kernels live in
// ensemble_kernels.h.
#include <cstring>

#include "ensemble_kernels.h"

using namespace scatrs;

// The forward pass of both entry points: the nodes from each ensemble's distribution (d_kind,
// d_p0, d_p1, d_p2 on the Gauss-Legendre rule of n_radius nodes) or, with d_nodes, the caller's
// n_radius (r, W) pairs per ensemble.
static int ensemble_forward(scatrs_ctx *ctx, int64_t count, const int32_t *d_kind,
                            const double *d_p0, const double *d_p1, const double *d_p2,
                            const double *d_nodes, const double *d_wavelength, int layers,
                            const double *d_fractions, const double *d_m, int64_t n_radius,
                            int64_t n_angles, int64_t order, int normalized,
                            int allow_unvalidated, double *d_scalars, double *d_greek,
                            double *d_phase, int64_t *total_terms) {
    if (!ctx) return SCATRS_INVALID_INPUT;
    ctx->error.clear();
    if (total_terms) *total_terms = 0;
    int rc = ens_check_args(ctx, count, layers, n_radius, n_angles, d_greek != nullptr, order);
    if (rc) return rc;
    if (count == 0) return SCATRS_OK;
    if ((!d_nodes && (!d_kind || !d_p0 || !d_p1)) || !d_wavelength || !d_m || !d_scalars ||
        (layers > 1 && !d_fractions))
        return fail(ctx, SCATRS_INVALID_INPUT, "null input or output pointer");
    EnsSetup S;
    if ((rc = ens_setup(ctx, count, d_kind, d_p0, d_p1, d_p2, d_wavelength, layers, d_fractions, d_m,
                        n_radius, n_angles, d_greek != nullptr, order, allow_unvalidated,
                        ctx->angular_kernel, S, d_nodes)))
        return rc;
    if (total_terms) *total_terms = S.total_terms;
    hipStream_t s = ctx->stream;
    // The phase at the caller's cosines (scatrs_set_phase_cosines): the angular products once
    // more, at the distinct |mu| as the non-negative cosines of a symmetric set of 2 h; the
    // Greek coefficients keep the angle rule's products (computed only when asked for).
    const bool user_mu = d_phase && !ctx->phase_mu.empty();
    PhaseMu U;
    if (user_mu && (rc = phase_mu_setup(ctx, S, U))) return rc;
    // Chunks of ensembles bounded by the staging buffer (a, b per term: 32 B, 16 B in fp32).
    const int64_t max_terms = chunk_terms_limit(4, S.stage32 ? sizeof(float4) : 2 * sizeof(cx));
    int64_t e0 = 0;
    while (e0 < count) {
        int64_t chunk_terms = 0;
        const int64_t e1 = next_chunk(S, e0, count, n_radius, max_terms, &chunk_terms);
        const int64_t ne = e1 - e0;
        ChunkBufs B;
        if ((rc = forward_nodes(ctx, S, e0, ne, chunk_terms, n_radius, n_angles, B, nullptr, s)))
            return rc;
        // The angle rule's products feed the Greek coefficients and, without the caller's
        // cosines, the phase matrix.
        if ((d_greek || (d_phase && !user_mu)) &&
            (rc = forward_angular(ctx, S, e0, ne, n_radius, n_angles, B, s)))
            return rc;
        double *tab_u = nullptr, *fac_u = nullptr;
        if (user_mu) {
            if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kAccumMu,
                                            (size_t)ne * (4 * 2 * (size_t)U.h + 1) * sizeof(double)),
                                "tables")))
                return rc;
            tab_u = static_cast<double *>(ctx->buf[scatrs_ctx::kAccumMu]);
            fac_u = tab_u + (size_t)ne * 4 * 2 * U.h;
            if ((rc = angular_products(ctx, S, e0, ne, n_radius, U.h, 2 * U.h, U.table, U.tab16,
                                       B, tab_u, s)))
                return rc;
        }
        double *gfac = nullptr;
        if (d_greek) {
            if ((rc = check_hip(ctx, ensure(ctx, scatrs_ctx::kGreekFac, ne * sizeof(double)),
                                "alloc")))
                return rc;
            gfac = static_cast<double *>(ctx->buf[scatrs_ctx::kGreekFac]);
        }
        ens_finish<<<(unsigned)ne, 256, 0, s>>>(e0, n_radius, (int)n_angles, S.d_atab, B.vals,
                                                B.tab_out, S.d_dtab, order, normalized,
                                                d_scalars, gfac, user_mu ? nullptr : d_phase,
                                                fac_u);
        if (user_mu) {
            const int64_t cells = ne * 6 * (int64_t)U.n_mu;
            ens_phase_mu<<<(unsigned)((cells + 255) / 256), 256, 0, s>>>(
                e0, ne, 2 * U.h, U.idx, U.n_mu, tab_u, fac_u, d_phase);
        }
        if (d_greek) {
            const int64_t len = order + 1;
            const dim3 grid((unsigned)((ne + 15) / 16), (unsigned)((len + 63) / 64));
            ens_greek_mfma<<<grid, 256, 0, s>>>(e0, ne, (int)n_angles, S.d_atab, B.tab_out,
                                                S.d_dtab, len, gfac, normalized, d_greek);
        }
        if ((rc = check_hip(ctx, hipGetLastError(), "ensemble launch"))) return rc;
        e0 = e1;
    }
    return SCATRS_OK;
}

extern "C" int scatrs_ensemble_fixed(scatrs_ctx *ctx, int64_t count, const int32_t *d_kind,
                                     const double *d_p0, const double *d_p1, const double *d_p2,
                                     const double *d_wavelength, int layers,
                                     const double *d_fractions, const double *d_m,
                                     int64_t n_radius, int64_t n_angles, int64_t order,
                                     int normalized, int allow_unvalidated, double *d_scalars,
                                     double *d_greek, double *d_phase, int64_t *total_terms) {
    return ensemble_forward(ctx, count, d_kind, d_p0, d_p1, d_p2, nullptr, d_wavelength, layers,
                            d_fractions, d_m, n_radius, n_angles, order, normalized,
                            allow_unvalidated, d_scalars, d_greek, d_phase, total_terms);
}

extern "C" int scatrs_ensemble_nodes(scatrs_ctx *ctx, int64_t count, int64_t n_nodes,
                                     const double *d_nodes, const double *d_wavelength,
                                     int layers, const double *d_fractions, const double *d_m,
                                     int64_t n_angles, int64_t order, int normalized,
                                     int allow_unvalidated, double *d_scalars, double *d_greek,
                                     double *d_phase, int64_t *total_terms) {
    if (ctx && count > 0 && !d_nodes) {
        ctx->error.clear();
        return fail(ctx, SCATRS_INVALID_INPUT, "null input or output pointer");
    }
    return ensemble_forward(ctx, count, nullptr, nullptr, nullptr, nullptr, d_nodes, d_wavelength,
                            layers, d_fractions, d_m, n_nodes, n_angles, order, normalized,
                            allow_unvalidated, d_scalars, d_greek, d_phase, total_terms);
}

extern "C" int scatrs_size_message(int32_t kind, const double *p, char *buf, int64_t cap) {
    if (!p || !buf || cap < 1) return SCATRS_INVALID_INPUT;
    const std::string msg = dist_message(kind, p);
    const size_t n = std::min<size_t>(msg.size(), (size_t)cap - 1);
    std::memcpy(buf, msg.data(), n);
    buf[n] = '\0';
    return msg.empty() ? SCATRS_OK : SCATRS_INVALID_INPUT;
}
