// The ensemble path on the fixed discrete grid (scatrs_ensemble_fixed): per ensemble (one
// distribution at one wavelength), n_radius Gauss-Legendre nodes in ln r, a Mie solve per node,
// the phase matrix at the n_angles Gauss-Legendre cosines by the parity-split angular
// products, the quadrature-weighted reduction, and the Greek projection. The kernels live in
// ensemble_kernels.h.
#include <cstring>
#include <string>
#include <vector>

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
                            double *d_phase, int64_t *total_terms, int tan_dir = -1,
                            double tan_sign = 0.0) {
    if (!ctx) return SCATRS_INVALID_INPUT;
    ctx->error.clear();
    if (total_terms) *total_terms = 0;
    int rc = ens_check_args(ctx, count, layers, n_radius, n_angles, d_greek != nullptr, order);
    if (rc) return rc;
    if (tan_dir >= 0 && (tan_dir >= 3 * layers - 1 || (tan_sign != 1.0 && tan_sign != -1.0)))
        return fail(ctx, SCATRS_INVALID_INPUT,
                    "the tangent direction must be in [0, " + std::to_string(3 * layers - 1) +
                        ") and its sign +1 or -1");
    if (count == 0) return SCATRS_OK;
    if ((!d_nodes && (!d_kind || !d_p0 || !d_p1)) || !d_wavelength || !d_m || !d_scalars ||
        (layers > 1 && !d_fractions))
        return fail(ctx, SCATRS_INVALID_INPUT, "null input or output pointer");
    EnsSetup S;
    // The angle rule's products (and so its table) feed the Greek coefficients and, without the
    // caller's cosines, the phase matrix.
    const bool rule_pass = d_greek || (d_phase && ctx->phase_mu.empty());
    if ((rc = ens_setup(ctx, count, d_kind, d_p0, d_p1, d_p2, d_wavelength, layers, d_fractions, d_m,
                        n_radius, n_angles, d_greek != nullptr, order, allow_unvalidated,
                        ctx->angular_kernel, S, d_nodes, rule_pass)))
        return rc;
    if (total_terms) *total_terms = S.total_terms;
    S.a.tan_dir = tan_dir;
    S.a.tan_sign = tan_sign;
    hipStream_t s = ctx->stream;
    // The phase at the caller's cosines (scatrs_set_phase_cosines): the angular products once
    // more, at the distinct |mu| as the non-negative cosines of a symmetric set of 2 h; the
    // Greek coefficients keep the angle rule's products (computed only when asked for).
    const bool user_mu = d_phase && !ctx->phase_mu.empty();
    // Without angular products the node solves need stage nothing (small launches then solve
    // one node per wavefront).
    S.values_only = !rule_pass && !user_mu;
    S.forward_only = true;
    PhaseMu U;
    if (user_mu && (rc = phase_mu_setup(ctx, S, U))) return rc;
    // Chunks of ensembles bounded by the staging buffer (a, b per term: 32 B, 16 B in fp32).
    auto &rp = ctx->replay;
    const int64_t max_terms =
        rp.mode == 2 ? rp.max_terms
                     : chunk_terms_limit(ctx->tune[SCATRS_TUNE_CHUNK_SHARE],
                                         S.stage32 ? sizeof(float4) : 2 * sizeof(cx));
    if (rp.mode == 1) rp.max_terms = max_terms;
    int64_t e0 = 0;
    while (e0 < count) {
        int64_t chunk_terms = 0;
        const int64_t e1 = next_chunk(S, e0, count, n_radius, max_terms, &chunk_terms);
        const int64_t ne = e1 - e0;
        ChunkBufs B;
        if ((rc = forward_nodes(ctx, S, e0, ne, chunk_terms, n_radius, n_angles, B, nullptr, s)))
            return rc;
        if (rule_pass && (rc = forward_angular(ctx, S, e0, ne, n_radius, n_angles, B, s)))
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

// A captured forward ensemble call: the HIP graph of its kernels, which own the device buffers
// they use (taken from the context when the graph was made).
struct scatrs_graph {
    int device = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> bufs;                    // the context's buffers and angle-rule tables
    int *node_n_rec = nullptr;                   // the recorded N per node, then the guard flag
    int *guard = nullptr;
    int *host_guard = nullptr;                   // pinned: the guard read back
    std::string error;
};

namespace {

// Moves every device buffer of the context into the graph, so that later calls on the context
// cannot free or grow memory the graph's kernels use (the context allocates afresh).
void take_buffers(scatrs_ctx *ctx, scatrs_graph *g) {
    for (int i = 0; i < scatrs_ctx::kNumBufs; ++i) {
        if (ctx->buf[i]) g->bufs.push_back(ctx->buf[i]);
        ctx->buf[i] = nullptr;
        ctx->cap[i] = 0;
    }
    for (auto &r : ctx->angle_rules) {
        if (r.d_atab) g->bufs.push_back(r.d_atab);
        if (r.d_dtab) g->bufs.push_back(r.d_dtab);
    }
    ctx->angle_rules.clear();
    ctx->rule_nr = -1;
}

void free_graph(scatrs_graph *g) {
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    for (void *p : g->bufs) (void)hipFree(p);
    if (g->node_n_rec) (void)hipFree(g->node_n_rec);
    if (g->host_guard) (void)hipHostFree(g->host_guard);
    delete g;
}

}  // namespace

extern "C" int scatrs_ensemble_fixed_graph(scatrs_ctx *ctx, int64_t count, const int32_t *d_kind,
                                           const double *d_p0, const double *d_p1,
                                           const double *d_p2, const double *d_wavelength,
                                           int layers, const double *d_fractions,
                                           const double *d_m, int64_t n_radius,
                                           int64_t n_angles, int64_t order, int normalized,
                                           int allow_unvalidated, double *d_scalars,
                                           double *d_greek, double *d_phase,
                                           int64_t *total_terms, scatrs_graph **out) {
    if (!ctx || !out) return SCATRS_INVALID_INPUT;
    *out = nullptr;
    if (count < 1) {
        ctx->error.clear();
        return fail(ctx, SCATRS_INVALID_INPUT, "a captured call needs at least one ensemble");
    }
    if (ctx->profiling)
        return fail(ctx, SCATRS_INVALID_INPUT, "turn profiling off before capturing a call");
    auto &rp = ctx->replay;
    rp = scatrs_ctx::Replay{};
    // 1. The call as usual, recording what the host reads back.
    rp.mode = 1;
    int rc = scatrs_ensemble_fixed(ctx, count, d_kind, d_p0, d_p1, d_p2, d_wavelength, layers,
                                   d_fractions, d_m, n_radius, n_angles, order, normalized,
                                   allow_unvalidated, d_scalars, d_greek, d_phase, total_terms);
    scatrs_graph *g = new scatrs_graph();
    g->device = ctx->device;
    g->node_n_rec = rp.node_n_rec;
    g->guard = rp.guard;
    if (rc == SCATRS_OK) rc = check_hip(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (rc == SCATRS_OK)
        rc = check_hip(ctx, hipHostMalloc(&g->host_guard, sizeof(int), hipHostMallocDefault),
                       "alloc");
    // 2. The same call captured from the record: kernels only.
    hipStream_t s = ctx->stream;
    hipStream_t own = nullptr;
    if (rc == SCATRS_OK && !s) {
        // The null stream cannot be captured: capture on a stream of our own.
        rc = check_hip(ctx, hipStreamCreateWithFlags(&own, hipStreamNonBlocking), "stream");
        ctx->stream = own;
    }
    if (rc == SCATRS_OK) {
        rp.mode = 2;
        rp.next_plan = 0;
        rc = check_hip(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal),
                       "capture");
        if (rc == SCATRS_OK) {
            int rc2 = scatrs_ensemble_fixed(ctx, count, d_kind, d_p0, d_p1, d_p2, d_wavelength,
                                            layers, d_fractions, d_m, n_radius, n_angles, order,
                                            normalized, allow_unvalidated, d_scalars, d_greek,
                                            d_phase, nullptr);
            hipGraph_t graph = nullptr;
            const hipError_t e = hipStreamEndCapture(ctx->stream, &graph);
            g->graph = graph;
            rc = rc2 ? rc2 : check_hip(ctx, e, "capture");
        }
        if (rc == SCATRS_OK)
            rc = check_hip(ctx, hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0),
                           "graph instantiate");
    }
    if (own) {
        (void)hipStreamSynchronize(own);
        (void)hipStreamDestroy(own);
        ctx->stream = nullptr;
    }
    rp.mode = 0;
    rp.node_n_rec = nullptr;
    rp.guard = nullptr;
    if (rc != SCATRS_OK) {
        free_graph(g);
        return rc;
    }
    take_buffers(ctx, g);
    *out = g;
    return SCATRS_OK;
}

extern "C" int scatrs_graph_launch(scatrs_graph *g, void *stream) {
    if (!g || !g->exec) return SCATRS_INVALID_INPUT;
    if (hipSetDevice(g->device) != hipSuccess) return SCATRS_DEVICE_ERROR;
    const hipError_t e = hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream));
    if (e != hipSuccess) {
        g->error = std::string("hipGraphLaunch: ") + hipGetErrorString(e);
        return SCATRS_DEVICE_ERROR;
    }
    return SCATRS_OK;
}

extern "C" int scatrs_graph_status(scatrs_graph *g, void *stream) {
    if (!g) return SCATRS_INVALID_INPUT;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (hipMemcpyAsync(g->host_guard, g->guard, sizeof(int), hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
        g->error = "could not read the graph's status";
        return SCATRS_DEVICE_ERROR;
    }
    if (*g->host_guard) {
        g->error =
            "the inputs changed the number of terms of a radius node since the call was "
            "captured: the replay kept the captured truncations and its results are invalid; "
            "capture the call again";
        return SCATRS_INVALID_INPUT;
    }
    g->error.clear();
    return SCATRS_OK;
}

extern "C" const char *scatrs_graph_last_error(const scatrs_graph *g) {
    return g ? g->error.c_str() : "";
}

extern "C" int scatrs_graph_destroy(scatrs_graph *g) {
    if (!g) return SCATRS_OK;
    (void)hipSetDevice(g->device);
    (void)hipDeviceSynchronize();
    free_graph(g);
    return SCATRS_OK;
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

extern "C" int scatrs_gauss_legendre(int64_t n, double *nodes, double *weights) {
    if (n < 1 || n > 65536 || !nodes || !weights) return SCATRS_INVALID_INPUT;
    std::vector<double> x, w;
    gauss_legendre(n, x, w);
    std::memcpy(nodes, x.data(), n * sizeof(double));
    std::memcpy(weights, w.data(), n * sizeof(double));
    return SCATRS_OK;
}

extern "C" int scatrs_ensemble_nodes_tangent(scatrs_ctx *ctx, int64_t count, int64_t n_nodes,
                                             const double *d_nodes, const double *d_wavelength,
                                             int layers, const double *d_fractions,
                                             const double *d_m, int64_t n_angles, int64_t order,
                                             int direction, int sign, int allow_unvalidated,
                                             double *d_scalars, double *d_greek, double *d_phase,
                                             int64_t *total_terms) {
    if (ctx && count > 0 && !d_nodes) {
        ctx->error.clear();
        return fail(ctx, SCATRS_INVALID_INPUT, "null input or output pointer");
    }
    if (ctx && direction < 0) {
        ctx->error.clear();
        return fail(ctx, SCATRS_INVALID_INPUT, "the tangent direction must be >= 0");
    }
    return ensemble_forward(ctx, count, nullptr, nullptr, nullptr, nullptr, d_nodes, d_wavelength,
                            layers, d_fractions, d_m, n_nodes, n_angles, order, 0,
                            allow_unvalidated, d_scalars, d_greek, d_phase, total_terms,
                            direction, (double)sign);
}

extern "C" int scatrs_size_message(int32_t kind, const double *p, char *buf, int64_t cap) {
    if (!p || !buf || cap < 1) return SCATRS_INVALID_INPUT;
    const std::string msg = dist_message(kind, p);
    const size_t n = std::min<size_t>(msg.size(), (size_t)cap - 1);
    std::memcpy(buf, msg.data(), n);
    buf[n] = '\0';
    return msg.empty() ? SCATRS_OK : SCATRS_INVALID_INPUT;
}
