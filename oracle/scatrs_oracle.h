/* CPU FP64 oracle for the size-distribution ensemble path (homogeneous and layered spheres).
 *
 * TEST INFRASTRUCTURE ONLY. This library restates the reference's algorithms in plain C so
 * that the tests, the smoke check and the benchmark's CPU-baseline leg can check and time the
 * GPU library against it. Nothing in the product (scatrs_amd/, include/) may link, load or
 * call it. Every function cites the reference file:line it follows.
 *
 * Parity pins (see tests/test_oracle_golden.py and tests/test_oracle_invariants.py):
 *   orc_mie_solve / orc_efficiencies*     -> Arb golden vectors mie.jsonl, mie_derivatives.jsonl
 *   orc_amplitudes                         -> Arb mie.jsonl series records (S1, S2)
 *   orc_layered_solve                      -> Arb layered.jsonl, layered_derivatives.jsonl;
 *                                             one layer == sphere bitwise
 *   orc_gauss_legendre                     -> Arb gauss_legendre.jsonl
 *   orc_wigner_d                           -> Arb wigner_d.jsonl
 *   orc_greek_project (sphere Greek)       -> Arb mie_greek.jsonl
 *   orc_ensemble_fixed                     -> invariants (monodisperse limit, alpha1_0 = 1,
 *                                             alpha1_1 = 3g, 1/2 int F11 = 1, Rayleigh r^6 law,
 *                                             closed-form moments, finite differences)
 *   orc_ensemble_fixed_jacobian            -> central differences of orc_ensemble_fixed_ex
 *                                             (moving and fixed range, fixed N per node);
 *                                             per-sphere Jacobians pinned by mie_derivatives.jsonl
 */
#ifndef SCATRS_ORACLE_H
#define SCATRS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    double re, im;
} orc_cx;

/* N = ceil(x + 6 x^(1/3) + 3) (mie/mod.rs:122-124,138-145). */
int64_t orc_default_n_max(double x);

/* First index of a downward ratio recurrence (mie/kernel.rs:99-112), for f64 (52 bits). */
int64_t orc_start_index(int64_t n_top, double rho_abs);

/* a_n, b_n (n = 1..n, index n-1) and the absorption sum (x^2/2) Q_abs of a sphere
 * (mie/kernel.rs:240-559). With derivs != 0, da_dx, da_dm, db_dx, db_dm and abs_jac[3] too
 * (any of the derivative pointers may be NULL when derivs == 0). Returns 0. */
int orc_mie_solve(double x, orc_cx m, int64_t n, int derivs, orc_cx *a, orc_cx *b,
                  orc_cx *da_dx, orc_cx *da_dm, orc_cx *db_dx, orc_cx *db_dm,
                  double *abs_sum, double *abs_jac);

/* [Q_ext, Q_sca, Q_abs, Q_back, g] from the coefficients (efficiencies.rs:64-84,169-187). */
void orc_efficiencies(double x, int64_t n, const orc_cx *a, const orc_cx *b, double abs_sum,
                      double out[5]);

/* The Jacobian [d/dx, d/dRe m, d/dIm m][Q_ext, Q_sca, Q_abs, Q_back, g]
 * (efficiencies.rs:89-146,192-230); `outer` adds the 2/x^2 term of the outer size parameter. */
void orc_efficiencies_jacobian(double x, int64_t n, const orc_cx *a, const orc_cx *b,
                               const orc_cx *da_dx, const orc_cx *da_dm, const orc_cx *db_dx,
                               const orc_cx *db_dm, double abs_sum, const double abs_jac[3],
                               int outer, double out[15]);

/* The batch entry behind _scatrs.mie_efficiencies (scatrs-py/src/mie.rs:303-364): values
 * (5, len) and, if jac != NULL, the Jacobian (3, 5, len). m is interleaved (re, im). n_max <= 0
 * takes the default truncation. Runs on `threads` OpenMP threads (<= 0: all). Returns 0, or
 * -(i+1) for the first sphere i whose inputs are invalid. */
int orc_mie_efficiencies_batch(const double *x, const double *m, int64_t len, int64_t n_max,
                               double *values, double *jac, int threads);

/* S1, S2 at the cosines mu (amplitudes.rs:52-65, the sums of Coefficients::amplitudes). */
void orc_amplitudes(int64_t n, const orc_cx *a, const orc_cx *b, const double *mu, int64_t n_mu,
                    orc_cx *s1, orc_cx *s2);

/* n-point Gauss-Legendre rule on [-1, 1], nodes ascending (special/gauss_legendre.rs:210-216):
 * Newton on P_n in long double from the Tricomi start, mirrored to exact symmetry. */
void orc_gauss_legendre(int64_t n, double *nodes, double *weights);

/* d^l_{mn}(x), l = 0..lmax (special/wigner.rs:152-182), for the pairs the projection uses:
 * (0,0), (2,2), (2,-2), (0,2). Returns 0, or -1 for another pair. */
int orc_wigner_d(int m, int n, double x, int64_t lmax, double *out);

/* Greek coefficients [alpha1, alpha2, alpha3, alpha4, beta1, beta2] (6 x (order+1)) of phase
 * tables f11, f12, f33, f34 at the n_q GL nodes (mie/greek.rs:161-221), unnormalized. */
void orc_greek_project(int64_t n_q, const double *nodes, const double *weights,
                       const double *f11, const double *f12, const double *f33,
                       const double *f34, int64_t order, double *out);

/* A layered sphere of `layers` layers (x nondecreasing, the outer last): a_n, b_n and the
 * absorption sum (layered/kernel.rs:308-465; one layer is the sphere's solve). With derivs,
 * per-layer derivative groups: dgrp[(l*4 + j)*n + k] for j = da_dx, da_dm, db_dx, db_dm and
 * abs_jac[l*3 + p] (layered/kernel.rs:479-587). Returns 0. */
int orc_layered_solve(int layers, const double *x, const orc_cx *m, int64_t n, int derivs,
                      orc_cx *a, orc_cx *b, orc_cx *dgrp, double *abs_sum, double *abs_jac);

/* Size distributions (ensemble/distribution.rs:15-68,366-430). */
enum {
    ORC_LOGNORMAL = 0,      /* p = (r_g, sigma_g, -) */
    ORC_GAMMA = 1,          /* p = (r_eff, v_eff, -) */
    ORC_MODIFIED_GAMMA = 2, /* p = (alpha, b, gamma) */
    ORC_POWER_LAW = 3,      /* p = (alpha, r_min, r_max): r_min, r_max fixed */
    ORC_MONODISPERSE = 4,   /* p = (radius, -, -) */
    ORC_SEGMENT = 5         /* p = (r_0, r_1, tau): n ~ (1 - tau)(1 - s) + tau s on [r_0, r_1] */
};

/* The unnormalized density u(r) of parameters p[3] with the reference radius fixed, and
 * grad[3] = du/dp (0 for fixed or unused parameters). */
double orc_density(int kind, const double *p, double r, double grad[3]);

/* The fixed discrete path of one ensemble (one distribution at one wavelength):
 *   n_radius Gauss-Legendre nodes in t = ln r over [t_lo, t_hi] (orc_radius_range);
 *   per node the Mie (or layered) solution, its per-particle quantities
 *   [1, C_ext, C_sca, C_sca g, C_abs, C_back, r^2, r^3] (ensemble/engine.rs:1046-1056) and its
 *   phase tables pi r^2 (4/x^2) [F11, F12, F33, F34] at the n_angles GL cosines;
 *   quadrature weights W_i = w_i (t_hi - t_lo)/2 u(r_i) r_i (w_i r for a monodisperse law, every
 *   node at its radius); number averages S/Z;
 *   outputs assembled as engine.rs:2396-2497 does: scalars[11] in SCALARS order, the Greek
 *   coefficients 6 x (order+1) (normalized by the projected alpha1_0 if normalized), and the
 *   phase matrix 6 x n_angles [F11, F12, F22, F33, F34, F44] (divided by <C_sca> if normalized).
 * layers == 1 is a sphere (fractions unused). Returns 0, or a negative code for invalid input. */
int orc_radius_range(int kind, const double *p, double *t_lo, double *t_hi);
int orc_ensemble_fixed(int kind, const double *p, double lambda, int layers,
                       const double *fractions, const orc_cx *m, int64_t n_radius,
                       int64_t n_angles, int64_t order, int normalized, double *scalars,
                       double *greek, double *phase);

/* orc_ensemble_fixed with the ln r range given as range[2] = (t_lo, t_hi) instead of the
 * distribution's, and N per node given as n_nodes[n_radius] (either may be NULL): for finite
 * differences at a fixed range and fixed truncations. */
int orc_ensemble_fixed_ex(int kind, const double *p, double lambda, int layers,
                          const double *fractions, const orc_cx *m, int64_t n_radius,
                          int64_t n_angles, int64_t order, int normalized, const double *range,
                          const int64_t *n_nodes, double *scalars, double *greek,
                          double *phase);

/* orc_ensemble_fixed_ex with the phase matrix at the caller's cosines mu_user[n_mu] (each in
 * [-1, 1]; phase is then 6 x n_mu), the Greek coefficients still from the n_angles-point rule
 * (the reference's mu argument, python/scatrs/ensemble.py:550-562). mu_user NULL: as _ex. */
int orc_ensemble_fixed_mu(int kind, const double *p, double lambda, int layers,
                          const double *fractions, const orc_cx *m, int64_t n_radius,
                          int64_t n_angles, int64_t order, int normalized, const double *range,
                          const int64_t *n_nodes, const double *mu_user, int64_t n_mu,
                          double *scalars, double *greek, double *phase);

/* The outputs of orc_ensemble_fixed and their forward-mode derivatives along the P = 3 + 2L +
 * (L - 1) directions [p0, p1, p2, Re m_0, Im m_0, ..., Re m_{L-1}, Im m_{L-1}, fraction_0 ..
 * fraction_{L-2}]: d_scalars (P, 11), d_greek (P, 6, order+1) or NULL, d_phase (P, 6, n_angles)
 * or NULL. fixed_range = 1 holds the ln r range fixed (the reference's derivative definition);
 * 0 moves the nodes with the parameters. n_nodes (may be NULL) receives N per node. */
int orc_ensemble_fixed_jacobian(int kind, const double *p, double lambda, int layers,
                                const double *fractions, const orc_cx *m, int64_t n_radius,
                                int64_t n_angles, int64_t order, int normalized,
                                int fixed_range, double *scalars, double *greek, double *phase,
                                double *d_scalars, double *d_greek, double *d_phase,
                                int64_t *n_nodes);

/* Many ensembles (each its own distribution and wavelength, spheres of one index per
 * ensemble), in parallel over ensembles on `threads` OpenMP threads: the CPU baseline. */
int orc_ensemble_batch(int64_t count, const int *kind, const double *p0, const double *p1,
                       const double *p2,
                       const double *lambda, const double *m, int64_t n_radius,
                       int64_t n_angles, int64_t order, int normalized, double *scalars,
                       double *greek, double *phase, int threads);

#ifdef __cplusplus
}
#endif
#endif
