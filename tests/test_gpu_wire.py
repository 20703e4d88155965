"""The wire-format entry points (scatrs_amd._wire, the reference's ``_scatrs.ensemble_compute``
and ``ensemble_compute_batch``) on the GPU: values and every derivative column against the
oracle's forward-mode Jacobian (the reference's derivative definition, range held fixed),
mixture weights against central differences, the batch's row order, and the two-rule error
estimates, ``converged`` and ``panels``."""

from __future__ import annotations

import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

OPTS = (1e-6, 4096, False, False)  # rtol, max_panels, allow_unvalidated, refine_backscatter


def _rel(got, ref, scale, tol, what):
    err = np.abs(got - ref) / np.where(scale > 0, scale, 1.0)
    assert err.max() <= tol, (what, err.max(), got, ref)


def test_single_lognormal_values_and_derivative_columns_match_the_oracle():
    """Values, Greek coefficients and the phase matrix at the caller's cosines, and d, greek_d,
    phase_d of all four parameters, against the oracle's closed-form forward Jacobian at a fixed
    range; the cosines are the oracle's Gauss-Legendre nodes, so its phase columns apply."""
    from scatrs_amd import _wire

    wl = np.array([0.45, 0.7, 1.1])
    m = np.array([1.45 + 2e-3j, 1.5 + 0.01j, 1.38 + 1e-5j])
    nr, na, L = 40, 40, 8
    mu, _ = oracle.gauss_legendre(na)
    items = [(("lognormal", 0.3, 1.6), ("sphere", m), 1.0)]
    names = _wire.ensemble_parameters(items, False)
    assert names == ["median_radius", "width", "m_real", "m_imag"]
    raw = _wire.ensemble_compute(items, wl, False, L, False, mu, True, [0, 1, 2, 3], *OPTS,
                                 n_radius=nr, n_angles=na)
    assert raw["values"].shape == (11, 3) and raw["d"].shape == (4, 11, 3)
    assert raw["greek"].shape == (6, 3, L + 1) and raw["greek_d"].shape == (4, 6, 3, L + 1)
    assert raw["phase"].shape == (6, 3, na) and raw["phase_d"].shape == (4, 6, 3, na)
    cols = [0, 1, 3, 4]  # the oracle's [p0, p1, p2, Re m, Im m]
    for i in range(3):
        J = oracle.ensemble_fixed_jacobian(0, 0.3, 1.6, wl[i], m[i], None, nr, na, L, True,
                                           True)
        s = J["scalars"]
        _rel(raw["values"][:, i], s, np.abs(s), 1e-10, ("values", i))
        _rel(raw["greek"][:, i], J["greek"], np.max(np.abs(J["greek"][0])), 1e-10, ("greek", i))
        _rel(raw["phase"][:, i], J["phase"], np.max(np.abs(J["phase"][0])), 1e-10, ("phase", i))
        ds = J["d_scalars"][cols]
        _rel(raw["d"][:, :, i], ds, np.max(np.abs(ds), 0), 1e-8, ("d", i))
        dg = J["d_greek"][cols]
        _rel(raw["greek_d"][:, :, i], dg, np.max(np.abs(dg), (1, 2))[:, None, None], 1e-8,
             ("greek_d", i))
        dp = J["d_phase"][cols]
        _rel(raw["phase_d"][:, :, i], dp, np.max(np.abs(dp), (1, 2))[:, None, None], 1e-8,
             ("phase_d", i))
    assert raw["converged"].dtype == bool and raw["panels"].dtype == np.uint64
    assert np.all(raw["panels"] == 1)


def test_layered_gamma_derivative_columns_match_the_oracle():
    """A gamma law of three-layer spheres: every layer's Re m, Im m and both radius fractions
    against the oracle's forward Jacobian (scalars and Greek coefficients)."""
    from scatrs_amd import _wire

    wl = np.array([0.5, 0.9])
    m = np.array([[1.6 + 0.1j, 1.7 + 0.05j], [1.45 + 1e-3j, 1.42 + 2e-3j],
                  [1.35 + 1e-5j, 1.33 + 1e-6j]])
    fr = np.array([0.4, 0.75])
    nr, na, L = 32, 48, 12
    items = [(("gamma", 0.4, 0.15), ("layered", fr, m), 1.0)]
    names = _wire.ensemble_parameters(items, False)
    raw = _wire.ensemble_compute(items, wl, False, L, False, np.zeros(0), True,
                                 list(range(len(names))), *OPTS, n_radius=nr, n_angles=na)
    assert "phase" not in raw
    for i in range(2):
        J = oracle.ensemble_fixed_jacobian(1, 0.4, 0.15, wl[i], m[:, i], fr, nr, na, L, True,
                                           True)
        cols = [0, 1] + list(range(3, J["d_scalars"].shape[0]))
        _rel(raw["values"][:, i], J["scalars"], np.abs(J["scalars"]), 1e-10, ("values", i))
        ds = J["d_scalars"][cols]
        _rel(raw["d"][:, :, i], ds, np.max(np.abs(ds), 0), 1e-8, ("d", i))
        dg = J["d_greek"][cols]
        _rel(raw["greek_d"][:, :, i], dg, np.max(np.abs(dg), (1, 2))[:, None, None], 1e-8,
             ("greek_d", i))


def test_mixture_values_and_weight_derivatives():
    """A mixture of a lognormal of spheres and a bimodal law (gamma + tabulated) of spheres:
    values against the oracle's mixture of the same parts, and the component and mode weight
    derivatives against central differences of the same GPU computation (h = 1e-5 relative:
    the outputs are rational in the weights)."""
    from scatrs_amd import _wire

    wl = np.array([0.55, 1.0])
    m0 = np.array([1.45 + 0.01j, 1.44 + 0.02j])
    m1 = np.array([1.33 + 1e-4j, 1.32 + 2e-4j])
    tab_r, tab_d = np.array([0.2, 0.5, 0.9]), np.array([1.0, 3.0, 0.5])

    def items(w0=2.0, w1=1.0, v0=3.0, v1=1.0):
        bi = ("multimodal", [(v0, ("gamma", 0.3, 0.1)), (v1, ("tabulated", tab_r, tab_d))])
        return [(("lognormal", 0.25, 1.5), ("sphere", m0), w0), (bi, ("sphere", m1), w1)]

    names = _wire.ensemble_parameters(items(), True)
    want = ["c0.weight", "c1.mode0.weight", "c1.mode1.weight", "c1.weight"]
    idx = [names.index(n) for n in want]
    kw = dict(n_radius=48, n_angles=48)
    L = 10
    raw = _wire.ensemble_compute(items(), wl, True, L, False, np.zeros(0), True, idx, *OPTS,
                                 **kw)
    # the parts' number fractions: weights 2 : 1, modes 3 : 1, segments by their integrals
    seg = np.array([0.5 * (1.0 + 3.0) * 0.3, 0.5 * (3.0 + 0.5) * 0.4])
    seg /= seg.sum()
    for i in range(2):
        parts = [(2 / 3, 0, 0.25, 1.5, m0[i], None),
                 (1 / 3 * 0.75, 1, 0.3, 0.1, m1[i], None),
                 (1 / 3 * 0.25 * seg[0], 5, 0.2, 0.5, m1[i], None, 3.0 / 4.0),
                 (1 / 3 * 0.25 * seg[1], 5, 0.5, 0.9, m1[i], None, 0.5 / 3.5)]
        ref = oracle.ensemble_mixture(parts, wl[i], order=L, **kw)
        _rel(raw["values"][:, i], ref["scalars"], np.abs(ref["scalars"]), 1e-10, ("values", i))
        _rel(raw["greek"][0, i], ref["greek"][0], np.max(np.abs(ref["greek"][0])), 1e-10,
             ("alpha1", i))
    assert np.all(raw["panels"] == 4)
    base = dict(w0=2.0, w1=1.0, v0=3.0, v1=1.0)
    for k, key in enumerate(["w0", "v0", "v1", "w1"]):
        h = 1e-5 * base[key]
        up = _wire.ensemble_compute(items(**dict(base, **{key: base[key] + h})), wl, True, L,
                                    False, np.zeros(0), True, [], *OPTS, **kw)
        dn = _wire.ensemble_compute(items(**dict(base, **{key: base[key] - h})), wl, True, L,
                                    False, np.zeros(0), True, [], *OPTS, **kw)
        fd = (up["values"] - dn["values"]) / (2 * h)
        got = raw["d"][k]
        scale = np.max(np.abs(fd), 1, keepdims=True)
        _rel(got, fd, scale, 1e-7, want[k])
        fdg = (up["greek"] - dn["greek"]) / (2 * h)
        _rel(raw["greek_d"][k], fdg, np.max(np.abs(fdg)), 1e-7, want[k] + " greek")


def test_batch_rows_are_distribution_major_and_match_single_calls():
    from scatrs_amd import _wire

    wl = np.array([0.4, 0.8, 1.6])
    m = np.array([1.5 + 1e-3j, 1.48 + 2e-3j, 1.46 + 5e-3j])
    specs = [("lognormal", 0.2, 1.7), ("gamma", 0.5, 0.2),
             ("multimodal", [(1.0, ("lognormal", 0.05, 1.3)), (0.2, ("power_law", 3.5, 0.1,
                                                                          2.0))])]
    mu = np.array([-1.0, -0.5, 0.0, 0.5, 1.0])
    kw = dict(n_radius=40, n_angles=64)
    b = _wire.ensemble_compute_batch(specs, ("sphere", m), wl, 16, False, mu, True, *OPTS, **kw)
    assert b["values"].shape == (11, 9) and b["phase"].shape == (6, 9, 5)
    assert b["d"].shape == (0, 11, 9)
    assert list(b["panels"]) == [1, 1, 1, 1, 1, 1, 2, 2, 2]
    for s, spec in enumerate(specs):
        one = _wire.ensemble_compute([(spec, ("sphere", m), 1.0)], wl, False, 16, False, mu,
                                     True, [], *OPTS, **kw)
        rows = slice(3 * s, 3 * s + 3)
        for key in ("values", "errors"):
            _rel(b[key][:, rows], one[key], np.abs(one["values"]), 1e-12, (key, s))
        _rel(b["greek"][:, rows], one["greek"], 1.0, 1e-12, ("greek", s))
        _rel(b["phase"][:, rows], one["phase"], np.max(np.abs(one["phase"])), 1e-12,
             ("phase", s))
        assert np.array_equal(b["converged"][rows], one["converged"])
    for i in range(3):
        ref = oracle.ensemble_fixed(1, 0.5, 0.2, wl[i], m[i], order=16, mu=mu, **kw)
        _rel(b["values"][:, 3 + i], ref["scalars"], np.abs(ref["scalars"]), 1e-10, i)


def test_errors_are_the_two_rule_difference_and_set_converged():
    """errors = |Q(n_radius) - Q(n_radius / 2)| per output (checked against the oracle at
    both rules); converged follows the reference's targets at rtol = 1e-6: a lognormal
    (r_g = 0.08, sigma_g = 1.6, m = 1.5 + 0.05i) on 128 nodes meets them at 1.2 um (the rules
    agree to 3e-9) but not at 0.55 um, where the sizes reach x = 9.6 and the rules differ by
    2.6e-5; on 8 nodes neither does."""
    from scatrs_amd import _wire

    wl = np.array([0.55, 1.2])
    m = np.array([1.5 + 0.05j, 1.5 + 0.05j])
    items = [(("lognormal", 0.08, 1.6), ("sphere", m), 1.0)]
    raw = _wire.ensemble_compute(items, wl, False, 8, False, np.zeros(0), True, [], *OPTS,
                                 n_radius=128, n_angles=64)
    for i in range(2):
        full = oracle.ensemble_fixed(0, 0.08, 1.6, wl[i], m[i], n_radius=128, n_angles=64,
                                     order=8)["scalars"]
        half = oracle.ensemble_fixed(0, 0.08, 1.6, wl[i], m[i], n_radius=64, n_angles=64,
                                     order=8)["scalars"]
        diff = np.abs(full - half)
        assert np.all(np.abs(raw["errors"][:, i] - diff) <= 1e-12 * np.abs(full) + 1e-300)
    assert list(raw["converged"]) == [False, True]
    coarse = _wire.ensemble_compute(items, wl, False, 8, False, np.zeros(0), True, [], *OPTS,
                                    n_radius=8, n_angles=64)
    assert not coarse["converged"].any()
    assert np.all(coarse["errors"][0] > 1e-6 * coarse["values"][0])


def test_every_greek_order_on_the_adaptive_engine():
    """all_orders (the reference's GreekOrder::All, engine.rs:2034-2037): each wavelength's
    coefficients up to twice the largest N of its final nodes, zero beyond (padded to the
    longest); exact projections, so alpha1_0 = 1 and alpha1_1 = 3 g; the coefficients equal
    the same call with that order given."""
    from scatrs_amd import _wire

    wl = np.array([0.45, 1.1])
    items = [(("lognormal", 0.1, 1.5), ("sphere", np.full(2, 1.5 + 0.01j)), 1.0)]
    raw = _wire.ensemble_compute(items, wl, False, None, True, np.zeros(0), True, [], *OPTS)
    a1 = raw["greek"][0]
    W = a1.shape[1]
    lengths = [int(np.max(np.flatnonzero(a1[i] != 0.0))) + 1 for i in range(2)]
    assert max(lengths) == W and lengths[0] != lengths[1]
    for i in range(2):  # orders 0 .. 2 N, zero beyond
        assert (lengths[i] - 1) % 2 == 0
        assert np.all(raw["greek"][:, i, lengths[i]:] == 0.0)
    assert np.allclose(a1[:, 0], 1.0, rtol=0, atol=1e-13)
    assert np.allclose(a1[:, 1], 3 * raw["values"][5], rtol=1e-10, atol=0)
    ref = _wire.ensemble_compute(items, wl, False, W - 1, False, np.zeros(0), True, [], *OPTS)
    for i in range(2):
        k = lengths[i]
        assert np.allclose(raw["greek"][:, i, :k], ref["greek"][:, i, :k], rtol=0, atol=1e-13)


def test_adaptive_errors_follow_the_references_propagation():
    """The adaptive engine's errors are the reference's bounds (lin.rs:24-44): the sums' errors
    add |Kronrod - Gauss| over the rows, and a ratio a/b takes (e_a + |a/b| e_b)/|b| (ssa =
    C_sca/C_ext, lidar ratio = 4 pi C_ext/C_back); the refined quantities' bounds meet rtol."""
    from scatrs_amd import _wire

    wl = np.array([0.5, 0.9])
    items = [(("lognormal", 0.3, 1.7), ("sphere", np.full(2, 1.45 + 2e-3j)), 1.0)]
    raw = _wire.ensemble_compute(items, wl, False, 6, False, np.zeros(0), True, [], 1e-6, 4096,
                                 False, False)
    v, e = raw["values"], raw["errors"]
    assert np.all(e[0] <= 1e-6 * v[0]) and np.all(e[1] <= 1e-6 * v[1])
    assert np.all(e[[0, 1, 2, 3]] > 0.0)
    assert np.allclose(e[4], (e[1] + v[4] * e[0]) / v[0], rtol=1e-12, atol=0)
    assert np.allclose(e[6], 4 * np.pi * (e[0] + v[6] / (4 * np.pi) * e[3]) / v[3], rtol=1e-12,
                       atol=0)
    assert np.all(e[8:11] == 0.0) and np.all(e[7] == 0.0)


def test_adaptive_mixture_of_layer_counts_with_derivatives():
    """A mixture of coated spheres (gamma law) and spheres (lognormal) on the adaptive engine,
    rows of both layer counts in each call: values equal Mixture.compute's, and the coated
    component's shell Im m and the lognormal's width derivatives match central differences
    of the engine (rtol 1e-10)."""
    from scatrs_amd import _wire

    wl = np.array([0.6])

    def items(mi=1e-3, width=1.6):
        coated = ("layered", np.array([0.6]), np.array([[1.6 + 0.1j], [1.4 + 1j * mi]]))
        return [(("gamma", 0.3, 0.12), coated, 1.0),
                (("lognormal", 0.25, width), ("sphere", np.array([1.5 + 0.01j])), 2.0)]

    names = _wire.ensemble_parameters(items(), True)
    want = ["c0.layer1.m_imag", "c1.width"]
    run = lambda its, d: _wire.ensemble_compute(its, wl, True, 6, False, np.zeros(0),  # noqa
                                                True, d, 1e-10, 4096, False, False)
    raw = run(items(), [names.index(n) for n in want])
    assert raw["converged"].all()
    for k, (key, base, h) in enumerate([("mi", 1e-3, 1e-6), ("width", 1.6, 1.6e-4)]):
        up = run(items(**{key: base + h}), [])
        dn = run(items(**{key: base - h}), [])
        fd = (up["values"] - dn["values"]) / (2 * h)
        sel = [0, 1, 2, 4, 5]
        scale = np.max(np.abs(fd[sel]), 1, keepdims=True)
        err = np.abs(raw["d"][k][sel] - fd[sel]) / scale
        assert err.max() <= 2e-5, (want[k], err.max())
        fdg = (up["greek"] - dn["greek"]) / (2 * h)
        assert np.max(np.abs(raw["greek_d"][k] - fdg)) <= 2e-5 * np.max(np.abs(fdg))


def test_adaptive_path_reports_sizes_outside_the_validated_domain():
    """A law whose range reaches x > 1e4 (layered/domain and mie/domain limits) raises
    OutsideValidatedDomainError with the reference's wavelength-and-radius context."""
    from scatrs_amd import _wire
    from scatrs_amd.errors import OutsideValidatedDomainError

    items = [(("lognormal", 200.0, 2.0), ("sphere", np.array([1.5 + 0.01j])), 1.0)]
    with pytest.raises(OutsideValidatedDomainError, match="at wavelength"):
        _wire.ensemble_compute(items, np.array([0.3]), False, None, False, np.zeros(0), True,
                               [], 1e-6, 4096, False, False)


def test_adaptive_unnormalized_outputs():
    """normalized=False on the adaptive engine (the reference's Normalization::Unnormalized):
    the same scalars; the Greek coefficients and the phase matrix times <C_sca>, so that
    alpha1_0 = <C_sca> (the projection is exact)."""
    from scatrs_amd import _wire

    wl = np.array([0.6, 1.0])
    items = [(("gamma", 0.4, 0.15), ("sphere", np.full(2, 1.45 + 0.01j)), 1.0)]
    mu = np.array([-0.5, 0.0, 0.7])
    a = _wire.ensemble_compute(items, wl, False, 8, False, mu, True, [], 1e-7, 4096, False,
                               False)
    b = _wire.ensemble_compute(items, wl, False, 8, False, mu, False, [], 1e-7, 4096, False,
                               False)
    assert np.array_equal(a["panels"], b["panels"])
    assert np.allclose(a["values"], b["values"], rtol=1e-14, atol=0)
    sca = a["values"][1]
    assert np.allclose(b["greek"][0, :, 0], sca, rtol=1e-12, atol=0)
    assert np.allclose(b["greek"], a["greek"] * sca[None, :, None], rtol=1e-12, atol=1e-300)
    assert np.allclose(b["phase"], a["phase"] * sca[None, :, None], rtol=1e-12, atol=0)
