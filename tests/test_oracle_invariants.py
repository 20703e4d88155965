"""The oracle's ensemble path against closed forms and invariants that pin it where no golden
vector exists (scatrs/tests/ensemble.rs:61-202 are the reference's own such checks)."""

from __future__ import annotations

import numpy as np
import pytest

import oracle


def test_moments_match_closed_forms():
    """<r^2>, <r^3> of a lognormal truncated to ln r_g -+ 5 ln sigma_g: r_g^k exp(k^2 s^2 / 2)
    [Phi(5 - k s) - Phi(-5 - k s)] / [Phi(5) - Phi(-5)] with s = ln sigma_g (the untruncated
    closed form is distribution.rs:221-233)."""
    from scipy.special import ndtr

    for rg, sg in ((0.3, 1.4), (0.08, 1.6), (1.0, 2.5)):
        r = oracle.ensemble_fixed(0, rg, sg, 0.5, 1.5 + 0j, n_radius=256, n_angles=8, order=-1)
        s = np.log(sg)
        mom = lambda k: rg ** k * np.exp(0.5 * k * k * s * s) * (ndtr(5 - k * s) - ndtr(-5 - k * s))
        z = ndtr(5.0) - ndtr(-5.0)
        assert abs(r["scalars"][9] / (np.pi * mom(2) / z) - 1) < 1e-12
        assert abs(r["scalars"][10] / (4 * np.pi / 3 * mom(3) / z) - 1) < 1e-12
        assert abs(r["scalars"][8] / (mom(3) / mom(2)) - 1) < 1e-12


def test_monodisperse_limit_is_the_sphere():
    """A narrow lognormal tends to the sphere (scatrs/tests/ensemble.rs:61-112)."""
    rg, wl, m = 0.4, 0.55, 1.5 + 0.01j
    r = oracle.ensemble_fixed(0, rg, 1.0 + 1e-5, wl, m, n_radius=32, n_angles=64, order=16)
    x = 2 * np.pi * rg / wl
    q = oracle.efficiencies(x, oracle.mie_solve(x, m))
    area = np.pi * rg * rg
    assert abs(r["scalars"][0] / (area * q[0]) - 1) < 1e-6
    assert abs(r["scalars"][5] / q[4] - 1) < 1e-6


def test_rayleigh_limit_scales_as_r6():
    """Tiny spheres: <C_sca> proportional to <r^6> (scatrs/tests/ensemble.rs:158-174)."""
    wl, m = 10.0, 1.5 + 0j
    a = oracle.ensemble_fixed(0, 0.001, 1.5, wl, m, n_radius=64, n_angles=8, order=-1)
    b = oracle.ensemble_fixed(0, 0.002, 1.5, wl, m, n_radius=64, n_angles=8, order=-1)
    assert abs(b["scalars"][1] / a["scalars"][1] / 64.0 - 1) < 1e-5


def test_phase_integral_and_greek_moments():
    r = oracle.ensemble_fixed(0, 0.2, 1.6, 0.5, 1.43 + 1e-8j, n_radius=96, n_angles=128,
                              order=32)
    mu, w = oracle.gauss_legendre(128)
    assert abs(0.5 * np.sum(w * r["phase"][0]) - 1) < 1e-13
    assert r["greek"][0, 0] == 1.0
    assert abs(r["greek"][0, 1] - 3 * r["scalars"][5]) < 1e-12
    # Greek of the Rayleigh limit: alpha1 = (1, 0, 1/2) (scatrs/tests/mie_greek.rs:476-496)
    s = oracle.ensemble_fixed(0, 0.0005, 1.2, 5.0, 1.5 + 0j, n_radius=32, n_angles=16, order=4)
    assert np.allclose(s["greek"][0, :3], [1.0, 0.0, 0.5], atol=1e-6)


def test_ensemble_derivative_of_the_discrete_path_by_finite_differences():
    """The fixed path is smooth in its parameters: central differences at h = 1e-6 relative
    agree between step sizes (a check of the oracle's smoothness the adjoint relies on)."""
    f = lambda rg: oracle.ensemble_fixed(0, rg, 1.5, 0.6, 1.5 + 0.01j, n_radius=64,
                                         n_angles=16, order=-1)["scalars"][1]
    rg = 0.25
    d1 = (f(rg * (1 + 1e-6)) - f(rg * (1 - 1e-6))) / (2e-6 * rg)
    d2 = (f(rg * (1 + 1e-5)) - f(rg * (1 - 1e-5))) / (2e-5 * rg)
    assert abs(d1 / d2 - 1) < 1e-6


def test_new_distributions_match_closed_forms():
    """The moments of the modified gamma law truncated to its range, in y = b r^gamma:
    <r^k> = b^(-k/gamma) Gamma(s_k) [P(s_k, y_hi) - P(s_k, y_lo)] / (Gamma(s_0) [P(s_0, y_hi) -
    P(s_0, y_lo)]), s_k = (alpha + 1 + k)/gamma (P the regularized incomplete gamma function;
    untruncated, the closed form of distribution.rs:234-238); the power law's on [r_min, r_max],
    int r^(k - alpha) / int r^(-alpha) (distribution.rs:239-252). A monodisperse law is the
    sphere (distribution.rs:49-53)."""
    from scipy.special import gammainc, gammaln

    for alpha, b, g in ((2.0, 3.0, 1.2), (-0.5, 8.0, 0.7), (6.0, 20.0, 2.0)):
        r = oracle.ensemble_fixed(oracle.MODIFIED_GAMMA, alpha, b, 0.5, 1.5 + 0j, n_radius=256,
                                  n_angles=8, order=-1, p2=g)
        t_lo, t_hi = oracle.radius_range(oracle.MODIFIED_GAMMA, alpha, b, g)
        y_lo, y_hi = b * np.exp(g * t_lo), b * np.exp(g * t_hi)

        def part(k):
            sk = (alpha + 1 + k) / g
            return np.exp(gammaln(sk)) * (gammainc(sk, y_hi) - gammainc(sk, y_lo))

        mom = lambda k: b ** (-k / g) * part(k) / part(0)  # noqa: E731
        assert abs(r["scalars"][9] / (np.pi * mom(2)) - 1) < 1e-11, (alpha, b, g)
        assert abs(r["scalars"][10] / (4 * np.pi / 3 * mom(3)) - 1) < 1e-11, (alpha, b, g)
        # the range holds all but e^-37 of the area-weighted law
        sa = (alpha + 3) / g
        assert gammainc(sa, y_lo) < 1e-16 and 1 - gammainc(sa, y_hi) < 1e-15
    for alpha, lo, hi in ((3.0, 0.05, 2.0), (-1.0, 0.1, 0.3), (2.0, 0.01, 1.0)):
        r = oracle.ensemble_fixed(oracle.POWER_LAW, alpha, lo, 0.5, 1.5 + 0j, n_radius=64,
                                  n_angles=8, order=-1, p2=hi)

        def integral(p):
            return np.log(hi / lo) if abs(p) < 1e-12 else (hi ** p - lo ** p) / p

        mom = lambda k: integral(k + 1 - alpha) / integral(1 - alpha)  # noqa: E731
        assert abs(r["scalars"][9] / (np.pi * mom(2)) - 1) < 1e-12, (alpha, lo, hi)
        assert abs(r["scalars"][8] / (mom(3) / mom(2)) - 1) < 1e-12, (alpha, lo, hi)
    rad, wl, m = 0.4, 0.55, 1.5 + 0.01j
    r = oracle.ensemble_fixed(oracle.MONODISPERSE, rad, 0.0, wl, m, n_radius=8, n_angles=64,
                              order=16)
    x = 2 * np.pi * rad / wl
    q = oracle.efficiencies(x, oracle.mie_solve(x, m))
    area = np.pi * rad * rad
    assert abs(r["scalars"][0] / (area * q[0]) - 1) < 1e-14
    assert abs(r["scalars"][5] / q[4] - 1) < 1e-14
    assert abs(r["scalars"][8] / rad - 1) < 1e-15
    with pytest.raises(ValueError):
        oracle.ensemble_fixed(oracle.MODIFIED_GAMMA, -1.0, 1.0, 0.5, 1.5 + 0j, p2=1.0)
    with pytest.raises(ValueError):
        oracle.ensemble_fixed(oracle.POWER_LAW, 2.0, 0.5, 0.5, 1.5 + 0j, p2=0.5)


def test_phase_at_the_callers_cosines():
    """The phase matrix at given cosines (the reference's mu argument): at the rule's own
    cosines it is the rule's phase matrix exactly; at mu = +-1 the sphere identities
    F12 = F34 = 0, F33 = F11 forward and -F11 backward; normalized, (1/2) sum w F11 = 1 over a
    second Gauss-Legendre rule; the scalars and Greek coefficients do not change."""
    kw = dict(n_radius=48, n_angles=64, order=16)
    args = (0, 0.2, 1.5, 0.55, 1.45 + 0.01j)
    a = oracle.ensemble_fixed(*args, **kw)
    mu64, _ = oracle.gauss_legendre(64)
    b = oracle.ensemble_fixed(*args, mu=mu64, **kw)
    assert np.array_equal(a["phase"], b["phase"]) and np.array_equal(a["greek"], b["greek"])
    assert np.array_equal(a["scalars"], b["scalars"])
    c = oracle.ensemble_fixed(*args, mu=[-1.0, 1.0], **kw)["phase"]
    f11 = c[0]
    assert np.all(np.abs(c[1]) <= 1e-15 * f11) and np.all(np.abs(c[4]) <= 1e-15 * f11)
    assert abs(c[3, 1] / f11[1] - 1) < 1e-15 and abs(c[3, 0] / f11[0] + 1) < 1e-15
    mu40, w40 = oracle.gauss_legendre(40)
    d = oracle.ensemble_fixed(*args, mu=mu40, **kw)["phase"]
    assert abs(0.5 * np.sum(w40 * d[0]) - 1) < 1e-12
