"""GPU parity of the batched-sphere path (scatrs_mie_efficiencies) against the oracle and the
reference's Arb golden vectors, through the C ABI."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.conftest import load_golden
from tests.tolerance import efficiency_conditions, efficiency_sizes, tolerance

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gpu(x, m, **kw):
    from scatrs_amd.mie import efficiencies

    return efficiencies(x, m, **kw)


def _stack(e):
    return np.stack([e.q_ext, e.q_sca, e.q_abs, e.q_back, e.g])


def test_arb_series_records_through_gpu():
    """Every (x, m) of the reference's Arb series records, solved on the GPU, within the
    reference's declared tolerance of the Arb value (mie/domain.rs:107-199)."""
    recs = [r for r in load_golden("mie.jsonl") if r["case"].endswith(",series")]
    x = np.array([r["inputs"]["x"] for r in recs])
    m = np.array([complex(r["inputs"]["m_re"], r["inputs"]["m_im"]) for r in recs])
    e = _gpu(x, m, derivatives=True)
    got = _stack(e)
    names = ["q_ext", "q_sca", "q_abs", "q_back", "g"]
    worst = 0.0
    for i, r in enumerate(recs):
        s = oracle.mie_solve(x[i], m[i], derivatives=True)
        jac = oracle.efficiencies_jacobian(x[i], s)
        q = got[:, i]
        sizes = efficiency_sizes(q, x[i])
        conds = efficiency_conditions(q, jac, x[i], m[i])
        for k, name in enumerate(names):
            ref = r["outputs"][name][0]
            if k == 2 and m[i].imag == 0.0:
                assert q[2] == 0.0
                continue
            err = abs(q[k] - ref) / sizes[k]
            tol = tolerance("efficiencies", x[i], m[i], conds[k])
            worst = max(worst, err / tol)
            assert err <= tol, (r["case"], name, q[k], ref, err, tol)
    print("worst GPU error / declared tolerance:", worst)


def _seeded(n, seed=7):
    from scatrs_amd import workloads

    return workloads.spheres(n, seed=seed)


def test_parity_with_oracle_on_batched_spheres():
    x, m = _seeded(20000)
    e = _gpu(x, m, derivatives=True)
    ref, jref = oracle.mie_efficiencies_batch(x, m, derivatives=True)
    got = _stack(e)
    for i in range(x.size):
        q = ref[:, i]
        sizes = efficiency_sizes(q, x[i])
        conds = efficiency_conditions(q, jref[:, :, i], x[i], m[i])
        for k in range(5):
            tol = 2.0 * tolerance("efficiencies", x[i], m[i], conds[k])
            assert abs(got[k, i] - q[k]) / sizes[k] <= tol, (i, k, x[i], m[i], got[k, i], q[k])
    # The Jacobian, in the reference's measure (mie/domain.rs:31-37): |error| over
    # max(|dQ/dp|, Q' (1 + |m| x) / |p|), Q' the size of Q in its value measure (Q_ext for
    # Q_abs), |p| = x for x and |m| for Re m, Im m.
    gj = e._jacobian
    worst = 0.0
    for i in range(x.size):
        sizes = efficiency_sizes(ref[:, i], x[i])
        sizes[2] = ref[0, i]
        for p in range(3):
            pmag = x[i] if p == 0 else abs(m[i])
            for k in range(5):
                floor = max(abs(jref[p, k, i]), sizes[k] * (1 + abs(m[i]) * x[i]) / pmag)
                worst = max(worst, abs(gj[p, k, i] - jref[p, k, i]) / floor)
    print("worst Jacobian error in the reference's measure:", worst)
    # The reference declares 256 (kappa max(1, |m|/|m-1|) + cond) 2^-52 for the Jacobian, cond
    # its own condition number (second derivatives, not available here); kappa alone reaches
    # ~1e-9 at x = 1000, and GPU and oracle each carry their own rounding, so 1e-7 bounds the
    # difference with room for cond.
    assert worst < 1e-7


@pytest.mark.parametrize("plain", [1, 0])
def test_values_without_derivatives_on_both_paths(plain):
    """The values-only batch on the one-sphere-per-lane kernel (SCATRS_TUNE_SPHERE_WAVE = 0; the
    branch-free plain path tried first, and with SCATRS_TUNE_SPHERE_PLAIN = 0 the checked path
    for every sphere) within 2x the reference's declared tolerance of the oracle, and the two
    paths within a few ulps of each other."""
    from scatrs_amd import _native, tune

    x, m = _seeded(20000, seed=11)
    ref, jref = oracle.mie_efficiencies_batch(x, m, derivatives=True)
    ctx = _native.context(0)
    try:
        ctx.set_tuning(tune.KEYS["sphere_wave"], 0)
        ctx.set_tuning(tune.KEYS["sphere_plain"], plain)
        assert ctx.get_tuning(tune.KEYS["sphere_plain"]) == plain
        got = _stack(_gpu(x, m))
        ctx.set_tuning(tune.KEYS["sphere_plain"], 1 - plain)
        other = _stack(_gpu(x, m))
    finally:
        ctx.set_tuning(tune.KEYS["sphere_plain"], 1)
        ctx.set_tuning(tune.KEYS["sphere_wave"], tune.DEFAULTS["sphere_wave"])
    for i in range(x.size):
        q = ref[:, i]
        sizes = efficiency_sizes(q, x[i])
        conds = efficiency_conditions(q, jref[:, :, i], x[i], m[i])
        for k in range(5):
            tol = 2.0 * tolerance("efficiencies", x[i], m[i], conds[k])
            assert abs(got[k, i] - q[k]) / sizes[k] <= tol, (i, k, x[i], m[i], got[k, i], q[k])
    # Q_ext, Q_sca and g of the two paths: the same recurrences, rounded differently only where
    # the plain path forms the two inverses of an order from one reciprocal.
    rel = np.abs(got - other)[[0, 1, 4]] / np.abs(other[[0, 1, 4]])
    assert rel.max() < 1e-12, rel.max()
    assert np.all(got[0] == got[1] + got[2])


def _with_tuning(key, value, fn):
    from scatrs_amd import _native, tune

    ctx = _native.context(0)
    old = ctx.get_tuning(tune.KEYS[key])
    ctx.set_tuning(tune.KEYS[key], value)
    try:
        return fn()
    finally:
        ctx.set_tuning(tune.KEYS[key], old)


def test_wavefront_per_sphere_path():
    """The values-only batch solved one sphere per wavefront (each sphere's recurrences split over
    the 64 lanes and joined by scans; SCATRS_TUNE_SPHERE_WAVE = 1 forces it) within 2x the
    reference's declared tolerance of the oracle and within rounding of the one-sphere-per-lane
    path, on the seeded batch and on spheres below x = 1, with a real index, and with N fixed
    below and above 64 terms."""
    x, m = _seeded(20000, seed=5)
    x = np.concatenate([x, [1e-3, 0.05, 0.7, 0.999, 1.0, 1.5, 2500.0, 9000.0]])
    m = np.concatenate([m, [1.5 + 0.1j, 1.33 + 0j, 2.0 + 1e-3j, 1.1 + 1e-6j, 1.5 + 0j,
                            1.33 + 1e-4j, 1.5 + 0.01j, 1.33 + 1e-8j]])
    ref, jref = oracle.mie_efficiencies_batch(x, m, derivatives=True)
    got = _stack(_with_tuning("sphere_wave", 1, lambda: _gpu(x, m)))
    lane = _stack(_with_tuning("sphere_wave", 0, lambda: _gpu(x, m)))
    worst = 0.0
    for i in range(x.size):
        q = ref[:, i]
        sizes = efficiency_sizes(q, x[i])
        conds = efficiency_conditions(q, jref[:, :, i], x[i], m[i])
        for k in range(5):
            tol = 2.0 * tolerance("efficiencies", x[i], m[i], conds[k])
            err = abs(got[k, i] - q[k]) / sizes[k]
            worst = max(worst, err / tol)
            assert err <= tol, (i, k, x[i], m[i], got[k, i], q[k])
    rel = np.abs(got - lane)[[0, 1, 4]] / np.abs(lane[[0, 1, 4]])
    print("worst error / 2x declared tolerance:", worst, "largest difference from the lane path:",
          rel.max())
    assert rel.max() < 1e-10, rel.max()
    assert np.all(got[0] == got[1] + got[2])
    for xs, ms, n in ((3.0, 1.5 + 0.01j, 20), (30.0, 1.33 + 1e-6j, 60), (150.0, 1.7 + 0.1j, 200)):
        w = _stack(_with_tuning("sphere_wave", 1,
                                lambda: _gpu(np.array([xs]), np.array([ms]), n_max=n)))
        s_ = oracle.mie_solve(xs, ms, n=n)
        assert np.allclose(w[:, 0], oracle.efficiencies(xs, s_), rtol=1e-11, atol=0)
    one = _stack(_with_tuning("sphere_wave", 1, lambda: _gpu(np.array([700.0]),
                                                             np.array([1.5 + 1e-3j]))))
    r1, _ = oracle.mie_efficiencies_batch(np.array([700.0]), np.array([1.5 + 1e-3j]))
    assert np.allclose(one[[0, 1, 4]], r1[[0, 1, 4]], rtol=1e-11, atol=0)


def test_wavefront_per_sphere_jacobian():
    """The Jacobian one sphere per wavefront (its pair terms across the lanes' boundaries
    included) against the one-per-lane kernel's and the oracle's, in the reference's measure."""
    x, m = _seeded(4000, seed=19)
    x = np.concatenate([x, [0.05, 0.7, 1.5, 2500.0]])
    m = np.concatenate([m, [1.33 + 0j, 2.0 + 1e-3j, 1.33 + 1e-4j, 1.5 + 0.01j]])
    w = _with_tuning("sphere_wave", 1, lambda: _gpu(x, m, derivatives=True))
    lane = _with_tuning("sphere_wave", 0, lambda: _gpu(x, m, derivatives=True))
    ref, jref = oracle.mie_efficiencies_batch(x, m, derivatives=True)
    gw, gl = w._jacobian, lane._jacobian
    worst_ref = worst_lane = 0.0
    for i in range(x.size):
        sizes = efficiency_sizes(ref[:, i], x[i])
        sizes[2] = ref[0, i]
        for p in range(3):
            pmag = x[i] if p == 0 else abs(m[i])
            for k in range(5):
                floor = max(abs(jref[p, k, i]), sizes[k] * (1 + abs(m[i]) * x[i]) / pmag)
                worst_ref = max(worst_ref, abs(gw[p, k, i] - jref[p, k, i]) / floor)
                worst_lane = max(worst_lane, abs(gw[p, k, i] - gl[p, k, i]) / floor)
    print("worst Jacobian error vs oracle:", worst_ref, "vs the lane kernel:", worst_lane)
    assert worst_ref < 1e-7
    assert worst_lane < 1e-9
    assert np.allclose(_stack(w), _stack(lane), rtol=1e-10, atol=0)


@pytest.mark.parametrize("n,derivatives", [(3000, False), (3000, True), (300_000, False)])
def test_host_array_entry_equals_the_device_entry(n, derivatives):
    """scatrs_mie_efficiencies_host (checks, N and ranking on the host, one copy each way)
    gives scatrs_mie_efficiencies' results bit for bit: a small batch (one sphere per
    wavefront), its Jacobian (one per lane) and a large batch (one per lane)."""
    from scatrs_amd.mie import efficiencies_device, efficiencies_host

    x, m = _seeded(n, seed=17)
    v, j, t = efficiencies_host(x, m, derivatives=derivatives)
    dev = torch.device("cuda", 0)
    vd, jd, td = efficiencies_device(torch.from_numpy(x).to(dev), torch.from_numpy(m).to(dev),
                                     derivatives=derivatives)
    assert t == td > 0
    assert np.array_equal(v, vd.cpu().numpy())
    if derivatives:
        assert np.array_equal(j, jd.cpu().numpy())
    from scatrs_amd import InvalidInputError

    xb = x[:10].copy()
    xb[7] = -1.0
    with pytest.raises(InvalidInputError, match="element 7: the size parameter"):
        efficiencies_host(xb, m[:10])


def test_real_index_gives_zero_absorption():
    x = np.array([0.5, 3.0, 40.0, 700.0])
    m = np.array([1.33, 1.5, 1.7, 1.45], dtype=np.complex128)
    e = _gpu(x, m)
    assert np.all(e.q_abs == 0.0)
    assert np.all(e.q_ext == e.q_sca)


def test_fixed_truncation_and_empty_input():
    x = np.array([5.0, 50.0])
    m = np.array([1.5 + 0.01j, 1.33 + 1e-6j])
    n = 200
    e = _gpu(x, m, n_max=n)
    for i in range(2):
        s = oracle.mie_solve(x[i], m[i], n=n)
        q = oracle.efficiencies(x[i], s)
        assert np.allclose(_stack(e)[:, i], q, rtol=1e-12, atol=0)
    e0 = _gpu(np.zeros(0), np.zeros(0, np.complex128))
    assert e0.q_ext.shape == (0,)


def test_errors_map_like_the_reference():
    from scatrs_amd import InvalidInputError, OutsideValidatedDomainError

    with pytest.raises(InvalidInputError, match="element 1: the size parameter"):
        _gpu(np.array([1.0, -2.0, 3.0]), np.array([1.5, 1.5, 1.5], np.complex128))
    with pytest.raises(InvalidInputError, match="gain medium"):
        _gpu(np.array([1.0]), np.array([1.5 - 0.1j]))
    with pytest.raises(OutsideValidatedDomainError, match="element 0"):
        _gpu(np.array([3e4]), np.array([1.5 + 0j]))
    e = _gpu(np.array([3e4]), np.array([1.5 + 0j]), allow_unvalidated=True)
    assert e.q_ext[0] > 1.9
    with pytest.raises(OutsideValidatedDomainError, match="below the default truncation"):
        _gpu(np.array([10.0]), np.array([1.5 + 0j]), n_max=5)


def test_full_size_properties():
    """At the benchmark's full size (2^20 spheres): bounded outputs everywhere, Q_ext =
    Q_sca + Q_abs, and a seeded subsample against the oracle."""
    from scatrs_amd.mie import efficiencies_device
    from scatrs_amd import workloads

    x, m = workloads.spheres(1 << 20)
    dev = torch.device("cuda", 0)
    values, _, terms = efficiencies_device(torch.from_numpy(x).to(dev),
                                           torch.from_numpy(m).to(dev))
    v = values.cpu().numpy()
    assert terms == workloads.mie_cost(x, m)[0]
    assert np.all(np.isfinite(v))
    assert np.all(v[2] >= 0) and np.all(v[1] > 0)
    assert np.all(np.abs(v[4]) <= 1.0)
    assert np.array_equal(v[0], v[1] + v[2])
    idx = np.random.default_rng(3).choice(x.size, 3000, replace=False)
    ref, _ = oracle.mie_efficiencies_batch(x[idx], m[idx])
    rel = np.abs(v[:, idx][[0, 1, 4]] - ref[[0, 1, 4]]) / np.abs(ref[[0, 1, 4]])
    assert rel.max() < 1e-9, rel.max()
