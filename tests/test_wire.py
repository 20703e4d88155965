"""The wire-format entry points (scatrs_amd._wire, the reference's ``_scatrs.ensemble_*``):
parameter names and every input check, all of which run before any device work."""

from __future__ import annotations

import numpy as np
import pytest

from scatrs_amd import _wire
from scatrs_amd.errors import InvalidInputError

M4 = np.full(4, 1.45 + 2e-3j)
WL4 = np.array([0.35, 0.5, 0.75, 1.0])
OPTS = dict(rtol=1e-6, max_panels=4096, allow_unvalidated=False, refine_backscatter=False)


def _compute(items, wl=WL4, mixture=False, greek=8, all_orders=False, mu=(), derivatives=(),
             **kw):
    o = dict(OPTS, **kw)
    return _wire.ensemble_compute(items, wl, mixture, greek, all_orders, np.asarray(mu, float),
                                  True, list(derivatives), o["rtol"], o["max_panels"],
                                  o["allow_unvalidated"], o["refine_backscatter"])


def test_parameter_names_follow_the_reference():
    """scatrs/src/ensemble/mod.rs:225-312: a multimodal law's modes then the particle's; in a
    mixture every name prefixed by its component and the component's weight last."""
    spec = ("multimodal", [(2.0, ("lognormal", 0.1, 1.5)), (1.0, ("gamma", 0.5, 0.1))])
    names = _wire.ensemble_parameters([(spec, ("sphere", M4), 1.0)], False)
    assert names == ["mode0.median_radius", "mode0.width", "mode0.weight",
                     "mode1.effective_radius", "mode1.effective_variance", "mode1.weight",
                     "m_real", "m_imag"]
    layered = ("layered", np.array([0.4, 0.7]), np.ones((3, 4), complex))
    items = [(("modified_gamma", 1.0, 2.0, 0.5), layered, 1.0),
             (("power_law", 3.0, 0.1, 1.0), ("sphere", M4), 2.0),
             (("tabulated", np.array([0.1, 0.2, 0.3]), np.array([1.0, 2.0, 0.0])),
              ("sphere", M4), 0.5),
             (("monodisperse", 0.3), ("sphere", M4), 0.5)]
    names = _wire.ensemble_parameters(items, True)
    assert names == ["c0.alpha", "c0.b", "c0.gamma", "c0.layer0.m_real", "c0.layer0.m_imag",
                     "c0.layer1.m_real", "c0.layer1.m_imag", "c0.layer2.m_real",
                     "c0.layer2.m_imag", "c0.layer0.radius_fraction",
                     "c0.layer1.radius_fraction", "c0.weight",
                     "c1.exponent", "c1.m_real", "c1.m_imag", "c1.weight",
                     "c2.m_real", "c2.m_imag", "c2.weight",
                     "c3.radius", "c3.m_real", "c3.m_imag", "c3.weight"]
    assert _wire.ensemble_parameters([], True) == []


def test_malformed_specs_raise_type_errors():
    with pytest.raises(TypeError, match="unknown distribution kind"):
        _wire.ensemble_parameters([(("weibull", 1.0, 2.0), ("sphere", M4), 1.0)], False)
    with pytest.raises(TypeError, match="must be a tuple"):
        _wire.ensemble_parameters([(["lognormal", 0.1, 1.5], ("sphere", M4), 1.0)], False)
    with pytest.raises(TypeError, match="unknown particle kind"):
        _wire.ensemble_parameters([(("lognormal", 0.1, 1.5), ("cube", M4), 1.0)], False)
    with pytest.raises(TypeError, match="must be an array"):
        _wire.ensemble_parameters([(("lognormal", 0.1, 1.5), ("sphere", 1.5), 1.0)], False)
    with pytest.raises(TypeError, match="shape \\(layers, wavelengths\\)"):
        _wire.ensemble_parameters(
            [(("lognormal", 0.1, 1.5), ("layered", np.array([0.5]), M4), 1.0)], False)
    with pytest.raises(InvalidInputError, match="not on the ensemble path"):
        _wire.ensemble_parameters([(("lognormal", 0.1, 1.5), ("spheroid", 2.0, M4), 1.0)],
                                  False)


def test_inputs_are_checked_as_the_reference_checks_them():
    one = [(("lognormal", 0.1, 1.5), ("sphere", M4), 1.0)]
    with pytest.raises(InvalidInputError, match="exactly one component, got 2"):
        _compute(one * 2)
    with pytest.raises(InvalidInputError, match="a mixture needs a component"):
        _compute([], mixture=True)
    with pytest.raises(InvalidInputError, match="no parameter 4"):
        _compute(one, derivatives=[4])
    with pytest.raises(InvalidInputError, match="parameter width requested twice"):
        _compute(one, derivatives=[1, 1])
    with pytest.raises(InvalidInputError, match="rtol must be in"):
        _compute(one, rtol=0.5)
    with pytest.raises(InvalidInputError, match="max_panels must be at least 1"):
        _compute(one, max_panels=0)
    with pytest.raises(InvalidInputError, match="at most 2097152"):
        _compute(one, greek=2 * (1 << 20) + 1)
    with pytest.raises(InvalidInputError, match="all_orders"):  # the fixed path's
        _wire.ensemble_compute(one, WL4, False, None, True, np.zeros(0), True, [], 1e-6, 4096,
                               False, False, n_radius=64)
    with pytest.raises(InvalidInputError, match="component 0 has 4 refractive indices for 3"):
        _compute(one, wl=WL4[:3])
    with pytest.raises(InvalidInputError, match="weight of component 1 must be finite"):
        _compute(one + [(("gamma", 0.5, 0.1), ("sphere", M4), -1.0)], mixture=True)
    with pytest.raises(InvalidInputError, match="finite, positive sum"):
        _compute([(("gamma", 0.5, 0.1), ("sphere", M4), 0.0)], mixture=True)
    with pytest.raises(InvalidInputError, match="MAX_MIE_VALUES"):
        _compute([(("lognormal", 0.1, 1.5), ("sphere", np.full(4000, 1.45 + 0j)), 1.0)],
                 wl=np.linspace(0.3, 1.0, 4000), greek=None, mu=np.linspace(-1, 1, 20000))
    with pytest.raises(InvalidInputError, match="radius fractions"):
        _wire.ensemble_compute_batch([("lognormal", 0.1, 1.5)],
                                     ("layered", np.array([0.2, 0.5]), np.ones((2, 4), complex)),
                                     WL4, 8, False, np.zeros(0), True, 1e-6, 4096, False, False)


def test_distribution_parameters_are_checked_before_any_device_work():
    """The reference's messages for invalid parameters (distribution.rs:99-185), from the
    library's host helper, for single laws, a multimodal law's mode and a batch."""
    with pytest.raises(InvalidInputError, match="the lognormal width must be finite and > 1"):
        _compute([(("lognormal", 0.1, 0.9), ("sphere", M4), 1.0)])
    with pytest.raises(InvalidInputError, match="effective_variance must be in"):
        _compute([(("gamma", 0.5, 0.7), ("sphere", M4), 1.0)])
    with pytest.raises(InvalidInputError, match="mode 1: alpha must be finite and > -1"):
        bi = ("multimodal", [(1.0, ("lognormal", 0.1, 1.5)),
                             (1.0, ("modified_gamma", -2.0, 1.0, 1.0))])
        _compute([(bi, ("sphere", M4), 1.0)])
    with pytest.raises(InvalidInputError, match="min_radius must be < max_radius"):
        _wire.ensemble_compute_batch([("lognormal", 0.1, 1.5), ("power_law", 3.0, 1.0, 0.5)],
                                     ("sphere", M4), WL4, None, False, np.zeros(0), True, 1e-6,
                                     4096, False, False)


def test_split_rules_pairs_the_gauss_copies_with_their_originals():
    """_split_rules on a doubled task list (T tasks, then their Gauss copies) with R replicas:
    panel rows [originals, copies] then monodisperse rows [originals, copies]; each half keeps
    its rows in order, the copies' keys mapped back to their originals', and its tasks' Z."""
    torch = pytest.importorskip("torch")
    T, R, nr, nm = 3, 2, 4, 1  # per half: 4 panel rows, 1 monodisperse row
    n_all = 2 * nr + 2 * nm
    place_task = [0, 1, 1, 2] + [3, 4, 4, 5] + [2] + [5]  # tasks of the rows, in place order
    key = torch.tensor([t * R + r for t in place_task for r in range(R)])
    lin = torch.arange(n_all * R * 7, dtype=torch.float64).reshape(n_all * R, 7)
    z = torch.arange(2 * T * R, dtype=torch.float64)
    rw = dict(key=key, lin=lin, gr=None, ph=None, z=z, n_rows=2 * nr, n_mono=2 * nm)
    avg = torch.arange(2 * T * R * 7, dtype=torch.float64).reshape(2 * T, R, 7)
    (ka, _, _, _, _, kd), (ga, _, _, _, _, gd) = _wire._split_rules(
        (avg, None, None, 0, 16, rw), T, R)
    assert torch.equal(ka, avg[:T]) and torch.equal(ga, avg[T:])
    assert torch.equal(kd["key"], gd["key"])  # the copies' keys are their originals'
    assert kd["lin"].shape == gd["lin"].shape == ((nr + nm) * R, 7)
    assert torch.equal(kd["lin"][:nr * R], lin[:nr * R])
    assert torch.equal(gd["lin"][:nr * R], lin[nr * R:2 * nr * R])
    assert torch.equal(gd["lin"][nr * R:], lin[(2 * nr + nm) * R:])
    assert torch.equal(kd["z"], z[:T * R]) and torch.equal(gd["z"], z[T * R:])


def test_results_from_the_wire_dict():
    """Ensemble results built from the wire format's dict: shapes of the grid and
    wavelengths, read-only arrays, error, d(name) with its own error, converged and panels."""
    from scatrs_amd.ensemble import _from_wire

    n, P, W, nmu = 4, 2, 3, 5
    rng = np.random.default_rng(0)
    raw = {"values": rng.random((11, n)), "errors": rng.random((11, n)),
           "d": rng.random((P, 11, n)), "d_errors": rng.random((P, 11, n)),
           "greek": rng.random((6, n, W)), "greek_errors": rng.random((6, n, W)),
           "greek_d": rng.random((P, 6, n, W)), "greek_d_errors": rng.random((P, 6, n, W)),
           "phase": rng.random((6, n, nmu)), "phase_errors": rng.random((6, n, nmu)),
           "phase_d": rng.random((P, 6, n, nmu)), "phase_d_errors": rng.random((P, 6, n, nmu)),
           "converged": np.array([True, False, True, True]),
           "panels": np.arange(n, dtype=np.uint64)}
    wl = np.array([[0.4, 0.5], [0.6, 0.7]])
    mu = np.linspace(-1, 1, nmu)
    r = _from_wire(raw, wl.shape, wl, mu, ["a", "b", "c"], [2, 0])
    assert r.params == ("c", "a") and r.c_ext.shape == (2, 2)
    assert np.array_equal(r.ssa.reshape(-1), raw["values"][4])
    assert np.array_equal(r.error.g.reshape(-1), raw["errors"][5])
    assert r.greek.alpha2.shape == (2, 2, W) and r.phase.f34.shape == (2, 2, nmu)
    assert np.array_equal(r.d("a").c_sca.reshape(-1), raw["d"][1][1])
    assert np.array_equal(r.d("c").error.greek.beta1.reshape(-1, W), raw["greek_d_errors"][0][4])
    assert np.array_equal(r.converged.reshape(-1), raw["converged"])
    assert not r.c_ext.flags.writeable
    with pytest.raises(KeyError):
        r.d("b")
