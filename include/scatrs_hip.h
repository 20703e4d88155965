/* scatrs_hip: MI355X (gfx950) library for the size-distribution ensemble path of scatrs over
 * homogeneous and layered spheres. C ABI: plain pointers and sizes, no C++ or torch types.
 *
 * Every pointer argument named d_* is DEVICE memory on the context's GPU; outputs are
 * caller-allocated. Calls are asynchronous on the context's stream except where a result
 * must come back to the host (validation status, sizing); each documents whether it
 * synchronizes. A context is not thread-safe: use one per host thread (the reference's
 * one-workspace-per-thread contract, scatrs/src/lib.rs:186-190).
 *
 * Status codes mirror the reference's error enum (scatrs/src/error.rs:14-32):
 * InvalidInput (nothing computed), OutsideValidatedDomain (inputs valid but outside the
 * validated domain, unless allow_unvalidated), NotConverged; plus a device-error code for
 * HIP failures. scatrs_last_error() returns the message of the last failure on a context,
 * with the element index or the wavelength and radius as the reference's messages carry them
 * (scatrs-py/src/mie.rs:108-117; scatrs/src/ensemble/engine.rs:124-131). */
#ifndef SCATRS_HIP_H
#define SCATRS_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCATRS_OK 0
#define SCATRS_INVALID_INPUT 1
#define SCATRS_OUTSIDE_VALIDATED_DOMAIN 2
#define SCATRS_NOT_CONVERGED 3
#define SCATRS_DEVICE_ERROR 4

#define SCATRS_ABI_VERSION 3

typedef struct scatrs_ctx scatrs_ctx;

/* ABI version of the loaded library (SCATRS_ABI_VERSION). */
int scatrs_abi_version(void);

/* Creates a context on HIP device `device`, issuing work on `stream` (a hipStream_t; NULL for
 * the device's null stream). No reference counterpart: the reference's workspaces are
 * `Mie::new()` / `Layered::new()` (scatrs/src/mie/mod.rs:190-198), which this replaces. */
int scatrs_create(int device, void *stream, scatrs_ctx **out);
int scatrs_destroy(scatrs_ctx *ctx);
int scatrs_set_stream(scatrs_ctx *ctx, void *stream);
const char *scatrs_last_error(const scatrs_ctx *ctx);

/* Instrumentation, no reference counterpart: with profiling on, the library brackets the
 * dominant kernels of each call with HIP events on the context's stream (up to 1024 regions):
 * a sphere batch, its persistent solve; a forward ensemble batch, per chunk of ensembles, the
 * node solves and then the angular products (two regions); a reverse ensemble batch, per chunk,
 * those two, then the reverse angular products (when Greek or phase cotangents are given) and
 * the reverse node solves.
 * scatrs_kernel_times waits for them, writes up to `cap` durations in milliseconds to the host
 * array ms_out, clears the record and returns how many it wrote. set_profiling clears it too. */
int scatrs_set_profiling(scatrs_ctx *ctx, int on);
int scatrs_kernel_times(scatrs_ctx *ctx, double *ms_out, int cap);

/* Selects the angular-product kernel of scatrs_ensemble_fixed and of the forward tables that
 * scatrs_ensemble_fixed_vjp recomputes: 0 (default) the FP64 matrix-core GEMM kernel, 1 the
 * plain FP64 vector kernel (same results to rounding; kept as the baseline the matrix-core
 * kernels are measured against), 2 the fp16 matrix cores with both operands split in two fp16
 * slices (exactly scaled per node and per table), three products summed in fp32 over 128
 * orders and then in fp64: the phase tables to about 1e-7 of their largest element, Greek
 * coefficients to about 1e-7 of max|alpha1| (the scalars do not depend on it); with n_radius
 * above 2048 kernel 2 falls back to kernel 0. The reverse products themselves always run in
 * FP64 (with kernel 2 the gradients stay within about 1e-7 of sum |cotangent x derivative|).
 * No reference counterpart. */
int scatrs_set_angular_kernel(scatrs_ctx *ctx, int kind);

/* Launch parameters of a context, for the project's autotuner (scatrs_amd/tune.py), which
 * times candidate values on the caller's workload and keeps the fastest; no reference
 * counterpart. Results do not depend on them, except SCATRS_TUNE_SPLIT_BLOCKS, which changes
 * the order in which a small batch's angular products are summed (rounding only).
 *   SCATRS_TUNE_SPHERE_WAVES: persistent wavefronts per CU of the sphere batch (1-32, default 16);
 *   SCATRS_TUNE_NODE_WAVES: the same for the ensembles' sphere node solves (1-32, default 8);
 *   SCATRS_TUNE_CHUNK_SHARE: a chunk of ensembles stages its coefficients in 1/share of the free
 *     device memory (1-64, default 4; at most 64 GiB, at least 1 GiB);
 *   SCATRS_TUNE_SPLIT_BLOCKS: blocks per CU that a small batch's FP64 matrix-core angular
 *     products aim for by dealing node groups over blocks (1-64, default 2);
 *   SCATRS_TUNE_SPHERE_PLAIN: 1 (default) solves each sphere of the sphere batch first on the
 *     branch-free plain path and redoes it on the checked path (the reference's branch per
 *     complex inverse) only where an argument left the plain range; 0 takes the checked path for
 *     every sphere (same results to rounding; for tests of that path);
 *   SCATRS_TUNE_SPHERE_WAVE: when a sphere batch (values and Jacobian; the layered entry points'
 *     one-layer case alike) is solved one sphere per wavefront
 *     (each sphere's recurrences split over the 64 lanes and joined by scans) instead of one per
 *     lane: 0 never, 1 always, f >= 2 (default 8) when its total terms times f are fewer than
 *     the lanes of the GPU times its largest N, i.e. when one sphere per lane would leave most
 *     lanes idle while the largest sphere's serial recurrences run (0-4096; same results to
 *     rounding).
 * scatrs_get_tuning returns the value, or -1 for an unknown key. */
#define SCATRS_TUNE_SPHERE_WAVES 0
#define SCATRS_TUNE_NODE_WAVES 1
#define SCATRS_TUNE_CHUNK_SHARE 2
#define SCATRS_TUNE_SPLIT_BLOCKS 3
#define SCATRS_TUNE_SPHERE_PLAIN 4
#define SCATRS_TUNE_SPHERE_WAVE 5
int scatrs_set_tuning(scatrs_ctx *ctx, int key, int64_t value);
int64_t scatrs_get_tuning(const scatrs_ctx *ctx, int key);

/* The cosines at which scatrs_ensemble_fixed writes the phase matrix: mu (host, n_mu values in
 * [-1, 1], at most 65536; n_mu = 0 restores the default, the n_angles Gauss-Legendre cosines);
 * d_phase is then (count, 6, n_mu). The angular products run once more at these cosines (with
 * the selected kernel); the Greek coefficients still come from the n_angles-point rule. The
 * reference's mu argument (python/scatrs/ensemble.py:550-562, scatrs-py/src/ensemble.rs:
 * 295-355). scatrs_ensemble_fixed_vjp then takes d_bar_phase (count, 6, n_mu) at these cosines
 * (its forward tables and reverse products run there too). Stays in effect until changed. */
int scatrs_set_phase_cosines(scatrs_ctx *ctx, const double *mu, int64_t n_mu);

/* The default truncation N = ceil(x + 6 x^(1/3) + 3) (host; scatrs/src/mie/mod.rs:122-124). */
int64_t scatrs_default_n_max(double x);

/* Efficiencies [Q_ext, Q_sca, Q_abs, Q_back, g] of `len` spheres, and with d_jacobian != NULL
 * their derivatives with respect to x, Re m and Im m. Replaces the batch loop of
 * _scatrs.mie_efficiencies (scatrs-py/src/mie.rs:303-364) over Mie::compute /
 * compute_with_derivatives (scatrs/src/mie/mod.rs:226-251) and MieSolution::efficiencies /
 * efficiencies_jacobian (scatrs/src/mie/efficiencies.rs:148-160).
 *   d_x[len], d_m[2*len] (re, im interleaved: complex128), n_max <= 0 for the default
 *   truncation; d_values (5, len) row-major; d_jacobian (3, 5, len) or NULL.
 *   total_terms (host, may be NULL): sum over spheres of N, the (particle, order) terms.
 * A batch too small to keep the GPU's lanes busy (SCATRS_TUNE_SPHERE_WAVE) is solved one sphere
 * per wavefront, its recurrences split over the lanes; larger batches one sphere per lane. Both
 * give the reference's results to rounding.
 * Synchronizes once (validation and sizing); the solve itself is queued on the stream. */
int scatrs_mie_efficiencies(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t len,
                            int64_t n_max, int allow_unvalidated, double *d_values,
                            double *d_jacobian, int64_t *total_terms);

/* scatrs_mie_efficiencies on host arrays, as _scatrs.mie_efficiencies takes them
 * (scatrs-py/src/mie.rs:303-364: numpy x and m in, numpy values and Jacobian out): x[len],
 * m[2*len], values (5, len), jacobian (3, 5, len) or NULL, all host memory. Up to 65536
 * spheres the checks, each sphere's N and the largest-first ranking are done on the host (the
 * first failing element reported as scatrs_mie_efficiencies reports it) and go to the device
 * with the inputs in one copy through the context's pinned staging: one synchronization, no
 * validation launches; larger batches are checked and ranked on the device. The same kernels
 * solve and the outputs come back in one copy. Results are those of scatrs_mie_efficiencies on
 * the same inputs. */
int scatrs_mie_efficiencies_host(scatrs_ctx *ctx, const double *x, const double *m, int64_t len,
                                 int64_t n_max, int allow_unvalidated, double *values,
                                 double *jacobian, int64_t *total_terms);

/* scatrs_mie_efficiencies without any device allocation: the same arguments, kernels and
 * results, with every buffer beyond the inputs and outputs (status, N histogram, keys,
 * permutation, the persistent wavefronts' slabs or, for a batch solved one sphere per
 * wavefront, their columns) carved from a 64 MiB array that the library
 * holds in each device's memory from its load until the process ends. No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The array's address on a device is
 * looked up when the first context on that device is created (scatrs_create).
 *   len <= 2^20 spheres, else SCATRS_INVALID_INPUT; spheres whose N needs more slab memory than
 *   the array holds for 4 wavefronts, SCATRS_INVALID_INPUT (about N > 20 000).
 * Every call on a device shares that device's array: calls that overlap on one device (two
 * streams, two threads) race and must be serialized by the caller. Synchronizes once, like
 * scatrs_mie_efficiencies. */
int scatrs_mie_efficiencies_noalloc(scatrs_ctx *ctx, const double *d_x, const double *d_m,
                                    int64_t len, int64_t n_max, int allow_unvalidated,
                                    double *d_values, double *d_jacobian, int64_t *total_terms);

/* The number of terms of each sphere without solving (host outputs, may be NULL): the largest
 * N, for sizing coefficient outputs, and the sum of N. The checks of Mie::compute; the
 * reference's mie::n_terms (scatrs/src/mie/mod.rs:126-145), as the sizing pass of
 * _scatrs.mie_coefficients uses it (scatrs-py/src/mie.rs:386-396). Synchronizes. */
int scatrs_mie_terms(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t len,
                     int64_t n_max, int allow_unvalidated, int64_t *max_terms,
                     int64_t *total_terms);

/* The Mie coefficients of `len` spheres: d_terms (len) each sphere's N; d_a, d_b (len, width)
 * complex128 a_n, b_n at [i, n - 1], zero beyond each sphere's N (width >= the largest N, from
 * scatrs_mie_terms, else SCATRS_INVALID_INPUT); with d_derivatives != NULL, (4, len, width)
 * complex128 [da/dx, da/dm, db/dx, db/dm] (the complex derivative d/dm; d/d(Im m) = i d/dm).
 * Replaces _scatrs.mie_coefficients (scatrs-py/src/mie.rs:376-442) over Mie::compute /
 * compute_with_derivatives and MieSolution::a, b, derivatives (scatrs/src/mie/mod.rs:226-251,
 * 356-380). Synchronizes once (validation). */
int scatrs_mie_coefficients(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t len,
                            int64_t n_max, int allow_unvalidated, int64_t width,
                            int64_t *d_terms, double *d_a, double *d_b, double *d_derivatives,
                            int64_t *total_terms);

/* Layered spheres: `count` particles of `layers` layers (1 <= layers <= 8), d_x (count, layers)
 * the interface size parameters core first (the outer one last, nondecreasing), d_m
 * (count, layers) complex128 the layers' indices, row-major like the PyO3 arrays
 * (scatrs-py/src/layered.rs:1-7). */

/* The largest and the total number of terms (host outputs, may be NULL) of layered spheres,
 * after the checks of Layered::compute (derivatives != 0: of compute_with_derivatives, whose
 * validated domain ends at x_L = 1e3); the reference's layered::n_terms
 * (scatrs/src/layered/mod.rs:217-226). Synchronizes. */
int scatrs_layered_terms(scatrs_ctx *ctx, const double *d_x, const double *d_m, int64_t count,
                         int layers, int64_t n_max, int allow_unvalidated, int derivatives,
                         int64_t *max_terms, int64_t *total_terms);

/* Efficiencies [Q_ext, Q_sca, Q_abs, Q_back, g] (5, count) of layered spheres per unit outer
 * cross section, and with d_jacobian != NULL their derivatives (layers, 3, 5, count) with
 * respect to each layer's x_l, Re m_l, Im m_l. Replaces _scatrs.layered_efficiencies
 * (scatrs-py/src/layered.rs:123-193) over Layered::compute / compute_with_derivatives,
 * efficiencies and efficiencies_jacobian (scatrs/src/layered/mod.rs:137-214,412-440).
 * Synchronizes once (validation). */
int scatrs_layered_efficiencies(scatrs_ctx *ctx, const double *d_x, const double *d_m,
                                int64_t count, int layers, int64_t n_max, int allow_unvalidated,
                                double *d_values, double *d_jacobian, int64_t *total_terms);

/* The Mie coefficients of layered spheres, as scatrs_mie_coefficients: d_terms (count), d_a,
 * d_b (count, width), and with d_derivatives != NULL (layers, 4, count, width)
 * [da/dx_l, da/dm_l, db/dx_l, db/dm_l] per layer. Replaces _scatrs.layered_coefficients
 * (scatrs-py/src/layered.rs:195-267) over LayeredSolution::a, b, derivatives
 * (scatrs/src/layered/mod.rs:341-380). Synchronizes once (validation). */
int scatrs_layered_coefficients(scatrs_ctx *ctx, const double *d_x, const double *d_m,
                                int64_t count, int layers, int64_t n_max, int allow_unvalidated,
                                int64_t width, int64_t *d_terms, double *d_a, double *d_b,
                                double *d_derivatives, int64_t *total_terms);

/* Size distributions (scatrs/src/ensemble/distribution.rs:15-68) and their parameters. */
#define SCATRS_LOGNORMAL 0      /* p0 = median radius r_g, p1 = width sigma_g */
#define SCATRS_GAMMA 1          /* p0 = effective radius, p1 = effective variance */
#define SCATRS_MODIFIED_GAMMA 2 /* n ~ r^alpha exp(-b r^gamma): p0 = alpha, p1 = b, p2 = gamma */
#define SCATRS_POWER_LAW 3      /* n ~ r^-alpha on [p1, p2]: p0 = alpha, p1 = r_min, p2 = r_max */
#define SCATRS_MONODISPERSE 4   /* p0 = radius */
/* One segment of a tabulated law: n linear on [p0, p1], n ~ (1 - p2)(1 - s) + p2 s,
 * s = (r - p0)/(p1 - p0), 0 <= p2 <= 1; no parameters to differentiate. A tabulated law
 * (distribution.rs:55-61) is the mixture of its segments weighted by their number integrals. */
#define SCATRS_SEGMENT 5

/* The fixed discrete path of a batch of ensembles (one distribution at one wavelength each;
 * the reference's EnsembleResult granularity, scatrs/src/ensemble/mod.rs:508-529):
 *   n_radius Gauss-Legendre nodes in ln r over the distribution's fixed range
 *   (lognormal: ln r_g +- 5 ln sigma_g; gamma: ln r_eff - 8 s .. ln r_eff +
 *   ln(1 + 37 v_eff + sqrt(74 v_eff)), s = sqrt(v_eff / (1 - 2 v_eff)); modified gamma, in
 *   y = b r^gamma: ln y from (ln Gamma(kA + 1) - 37)/kA to ln(kA + 37 + sqrt(74 kA)),
 *   kA = (alpha + 3)/gamma, t = (ln y - ln b)/gamma; power law: its
 *   support; monodisperse: every node at the radius, weights w r), weights
 *   w (t_hi - t_lo)/2 u(r) r;
 *   d_p0, d_p1 (count) and d_p2 (count; may be NULL when no ensemble needs it: 0) the
 *   distribution parameters as the SCATRS_* kinds list them;
 *   the phase matrix at the n_angles-point Gauss-Legendre cosines (2 <= n_angles <= 4096);
 *   Greek to `order` (d_greek == NULL: none).
 * Spheres: layers == 1 and d_m (count, 2) is each ensemble's index. Layered spheres:
 *   layers = L > 1, d_fractions (count, L-1) the interface radii as fractions of the outer
 *   radius (nondecreasing in (0, 1]), d_m (count, L, 2) each layer's index.
 * Outputs (normalized as the reference's Normalization::Normalized when normalized != 0):
 *   d_scalars (count, 11) in the reference's SCALARS order (python/scatrs/ensemble.py:51-63);
 *   d_greek (count, 6, order+1) [alpha1..alpha4, beta1, beta2] or NULL;
 *   d_phase (count, 6, n_angles) [F11, F12, F22, F33, F34, F44] or NULL ((count, 6, n_mu) at
 *   the cosines of scatrs_set_phase_cosines when set).
 * Replaces ensemble::compute_with_particle / compute_batch_with_particle
 * (scatrs/src/ensemble/mod.rs:589-686) and _scatrs.ensemble_compute{,_batch}
 * (scatrs-py/src/ensemble.rs:295-355,456-500) on this fixed discrete path (the reference's
 * adaptive Gauss-Kronrod panels are not part of this path). Synchronizes once per call. */
int scatrs_ensemble_fixed(scatrs_ctx *ctx, int64_t count, const int32_t *d_kind,
                          const double *d_p0, const double *d_p1, const double *d_p2,
                          const double *d_wavelength,
                          int layers, const double *d_fractions, const double *d_m,
                          int64_t n_radius, int64_t n_angles, int64_t order, int normalized,
                          int allow_unvalidated, double *d_scalars, double *d_greek,
                          double *d_phase, int64_t *total_terms);

/* Reverse mode of scatrs_ensemble_fixed (spheres or layered spheres): given the cotangents of
 * its outputs (d_bar_scalars (count, 11), d_bar_greek (count, 6, order+1), d_bar_phase
 * (count, 6, n_angles); each may be NULL for zero), the vector-Jacobian product with respect to
 * each ensemble's parameters, d_grad (count, P) with P = 3 + 2 L + (L - 1):
 *   [p0, p1, p2, Re m_0, Im m_0, ..., Re m_{L-1}, Im m_{L-1}, fraction_0 .. fraction_{L-2}]
 * (0 in the columns of unused parameters and of a power law's fixed limits, the reference's
 * parameter_names, distribution.rs:189-198)
 * (layer 0 the core; the fractions act at a fixed outer radius, the reference's
 * layer{l}.radius_fraction direction, scatrs/src/ensemble/particle.rs:221-249).
 * fixed_range = 0: the derivative of this discrete path, whose radius nodes and range move with
 * the parameters; fixed_range = 1: the ln r range held at its value, so the parameters act
 * through the density alone, the reference's definition (a monodisperse law's range is its
 * radius, which moves its nodes either way) (docs/theory/ensembles.md:150,339-340,
 * ensemble/engine.rs:1793-1804). a_n, b_n and their derivatives are recomputed, not stored.
 * Replaces the reference's forward-mode derivatives (EnsembleOptions::derivatives,
 * scatrs/src/ensemble/mod.rs:356-357, contracted by the caller) with one reverse pass. */
int scatrs_ensemble_fixed_vjp(scatrs_ctx *ctx, int64_t count, const int32_t *d_kind,
                              const double *d_p0, const double *d_p1, const double *d_p2,
                              const double *d_wavelength, int layers,
                              const double *d_fractions, const double *d_m, int64_t n_radius,
                              int64_t n_angles, int64_t order, int normalized, int fixed_range,
                              int allow_unvalidated, const double *d_bar_scalars,
                              const double *d_bar_greek, const double *d_bar_phase,
                              double *d_grad, int64_t *total_terms);

/* A forward ensemble call captured as a HIP graph, for replaying the same batch without
 * launch or host overhead (a latency-bound call such as three ensembles of 128 nodes issues
 * about twenty kernels and synchronizes twice). No reference counterpart: the reference has no
 * device. scatrs_ensemble_fixed_graph takes scatrs_ensemble_fixed's arguments, runs the call
 * once (the outputs are written, errors are reported as scatrs_ensemble_fixed's), then captures
 * the same call's kernels, with every host-side decision (sizes, chunks, launch plans) taken
 * from that run, into *out. The graph owns the device buffers the call used (the context
 * allocates afresh for later calls), so that other calls on the context cannot disturb it.
 * scatrs_graph_launch replays the kernels on `stream` (a hipStream_t; NULL: the null stream),
 * reading the input buffers and writing the output buffers of the capture: inputs changed in
 * place are recomputed as long as every radius node keeps its number of terms N; otherwise the
 * replay keeps the captured N (so that every launch stays within its memory) and
 * scatrs_graph_status (synchronizes `stream`) returns SCATRS_INVALID_INPUT with a message in
 * scatrs_graph_last_error. A graph is not thread-safe; replays on one graph must not overlap.
 * scatrs_graph_destroy frees it (synchronizes the device). */
typedef struct scatrs_graph scatrs_graph;
int scatrs_ensemble_fixed_graph(scatrs_ctx *ctx, int64_t count, const int32_t *d_kind,
                                const double *d_p0, const double *d_p1, const double *d_p2,
                                const double *d_wavelength, int layers,
                                const double *d_fractions, const double *d_m, int64_t n_radius,
                                int64_t n_angles, int64_t order, int normalized,
                                int allow_unvalidated, double *d_scalars, double *d_greek,
                                double *d_phase, int64_t *total_terms, scatrs_graph **out);
int scatrs_graph_launch(scatrs_graph *g, void *stream);
int scatrs_graph_status(scatrs_graph *g, void *stream);
const char *scatrs_graph_last_error(const scatrs_graph *g);
int scatrs_graph_destroy(scatrs_graph *g);

/* The ensembles of scatrs_ensemble_fixed on the caller's radius nodes: d_nodes (count, n_nodes,
 * 2) the (r, W) pairs of each ensemble's quadrature (r > 0, W >= 0; W the quadrature weight
 * times the number density per ln r, so that an average is sum W q / sum W), every other
 * argument and output as scatrs_ensemble_fixed's (n_nodes in place of n_radius). The node
 * evaluator of the reference's adaptive engine: pass 1 (engine.rs:307-943) evaluates the
 * Kronrod and Gauss nodes of its panels (kronrod.rs:5-52), pass 2 (engine.rs:1423-1766) all
 * the panels' nodes at once; the host places the panels (scatrs_amd/adaptive.py). */
int scatrs_ensemble_nodes(scatrs_ctx *ctx, int64_t count, int64_t n_nodes, const double *d_nodes,
                          const double *d_wavelength, int layers, const double *d_fractions,
                          const double *d_m, int64_t n_angles, int64_t order, int normalized,
                          int allow_unvalidated, double *d_scalars, double *d_greek,
                          double *d_phase, int64_t *total_terms);

/* Forward mode of scatrs_ensemble_nodes along one of the particle's directions (direction 2 l,
 * 2 l + 1: Re m, Im m of layer l, core first; 2 L + j: radius fraction j as x_L d/dx_j at a
 * fixed outer size): the same call (normalized = 0) with every node's coefficients moved to
 * c + sign eps dc/d(direction), eps = 1/(1 + x) at the node's outer size parameter x, and its
 * weight W scaled by 1/(2 eps). Every output's weighted sum is linear or quadratic in the
 * coefficients, so the sign = +1 call's sums (average times the row's scaled weight) minus the
 * sign = -1 call's are the sums' derivative along the direction, exactly up to rounding; the
 * absorption sum moves by sign eps times its derivative along the direction (the absorption
 * identity's Jacobian: the outer layer's from the solve, an inner layer's summed over the orders
 * from its sensitivities), so the C_abs and C_ext columns are exact too. Replaces the
 * reference's forward-mode derivative columns (ensemble/engine.rs:1111-1123, particle.rs:221-249)
 * for the particle's directions: P + 1 evaluations instead of one reverse pass per output. */
int scatrs_ensemble_nodes_tangent(scatrs_ctx *ctx, int64_t count, int64_t n_nodes,
                                  const double *d_nodes, const double *d_wavelength, int layers,
                                  const double *d_fractions, const double *d_m, int64_t n_angles,
                                  int64_t order, int direction, int sign, int allow_unvalidated,
                                  double *d_scalars, double *d_greek, double *d_phase,
                                  int64_t *total_terms);

/* Reverse mode of scatrs_ensemble_nodes, as scatrs_ensemble_fixed_vjp's with the range held
 * fixed: the nodes stay where they are, and the parameters act through each node's ln W, whose
 * derivatives d_dlnw (count, n_nodes, 3) the caller gives (d ln u(r) / d p at the node for
 * W = w u(r) r; may be NULL: columns 0-2 zero). d_grad (count, P) as scatrs_ensemble_fixed_vjp's:
 * columns 0-2 sum phi_i d ln W_i / d p over the nodes, phi_i the cotangent of ln W_i; then
 * Re m, Im m per layer and the radius fractions. The adaptive engine's derivatives
 * (engine.rs:1111-1123,1793-1804) on its final panels. */
int scatrs_ensemble_nodes_vjp(scatrs_ctx *ctx, int64_t count, int64_t n_nodes,
                              const double *d_nodes, const double *d_dlnw,
                              const double *d_wavelength, int layers, const double *d_fractions,
                              const double *d_m, int64_t n_angles, int64_t order, int normalized,
                              int allow_unvalidated, const double *d_bar_scalars,
                              const double *d_bar_greek, const double *d_bar_phase,
                              double *d_grad, int64_t *total_terms);

/* Host helpers of the adaptive engine, with the kernels' definitions: the fixed ln r range of
 * a distribution (p[3] as the SCATRS_* kinds list them), and ln u(r) of the number density per
 * unit radius, up to a constant, at n points, each with its own distribution (kind[i],
 * p[3 i .. 3 i + 2]) and radius r[i] (distribution.rs:366-430; the density per ln r is u r),
 * with d_ln_u (n, 3) not NULL also d ln u / d p (the reference radius held fixed).
 * SCATRS_INVALID_INPUT for invalid parameters. No context, no device work. A log-density
 * call of more than 32768 points runs on up to 16 host threads that it starts and joins
 * before it returns. The results do not depend on the thread count. */
int scatrs_size_range(int32_t kind, const double *p, double *t_lo, double *t_hi);
/* The reference's message for invalid parameters p[3] of a distribution (distribution.rs:
 * 99-185) into buf (cap bytes, NUL-terminated): SCATRS_INVALID_INPUT with the message, or
 * SCATRS_OK and "" for valid ones. */
int scatrs_size_message(int32_t kind, const double *p, char *buf, int64_t cap);
int scatrs_size_log_density(int64_t n, const int32_t *kind, const double *p, const double *r,
                            double *ln_u, double *d_ln_u);
/* The n-point Gauss-Legendre rule on [-1, 1] the kernels use (radius nodes in ln r, the
 * angular rule, and the cosines of the phase matrix at n_angles), nodes ascending and exactly
 * symmetric, into host arrays of n doubles; 1 <= n <= 65536, the largest rule the library
 * builds. Replaces scatrs.special.gauss_legendre (python/scatrs/special.py:31-68,
 * scatrs/src/special/gauss_legendre.rs:210-216) for the rules of this path. Host only. */
int scatrs_gauss_legendre(int64_t n, double *nodes, double *weights);

#ifdef __cplusplus
}
#endif
#endif
