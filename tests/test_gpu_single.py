"""The single-particle entry points through the C ABI on the GPU: sphere coefficients
(scatrs_mie_terms, scatrs_mie_coefficients) against the reference's Arb records in mie.jsonl and
mie_derivatives.jsonl, and layered spheres (scatrs_layered_terms, scatrs_layered_coefficients,
scatrs_layered_efficiencies) against layered.jsonl and layered_derivatives.jsonl, each under the
reference's declared tolerance model, plus the oracle and the size-independent identities."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.conftest import load_golden
from tests.golden_checks import (check_coefficients, check_efficiencies, coefficient_n, cvals,
                                 layered_jacobian, layered_records, rec_inputs)
from tests.tolerance import efficiency_sizes, tolerance

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _m(rec):
    i = rec["inputs"]
    return complex(i["m_re"], i["m_im"])


def test_mie_coefficients_match_arb():
    """Every coefficient record of mie.jsonl in one batch: a_n, b_n within the declared
    tolerance (mie/domain.rs:107-199), n_terms the default truncation, zero padding."""
    from scatrs_amd.mie import coefficients, default_n_max

    recs = [r for r in load_golden("mie.jsonl") if r["case"].endswith(",coefficients")]
    x = np.array([r["inputs"]["x"] for r in recs])
    m = np.array([_m(r) for r in recs])
    # the records store orders up to the default truncation
    c = coefficients(x, m, derivatives=True)
    assert np.array_equal(c.n_terms, default_n_max(x))
    assert c.a.shape == (x.size, c.n_terms.max())
    worst = 0.0
    for i, r in enumerate(recs):
        n = c.n_terms[i]
        assert np.all(c.a[i, n:] == 0) and np.all(c.b[i, n:] == 0)
        orders = np.array(r["inputs"]["orders"])
        dx, dm = c.d("x"), c.d("m")
        for stem, got, gx, gm in (("a", c.a[i], dx.a[i], dm.a[i]), ("b", c.b[i], dx.b[i],
                                                                     dm.b[i])):
            ref = cvals(r, stem)
            scale = np.max(np.abs(ref))
            err = np.max(np.abs(got[orders - 1] - ref)) / scale
            cond = np.max(x[i] * np.abs(gx[orders - 1]) + abs(m[i]) * np.abs(gm[orders - 1])) / scale
            tol = tolerance("coefficients", x[i], m[i], cond)
            worst = max(worst, err / tol)
            assert err <= tol, (r["case"], stem, err, tol)
    print("worst GPU coefficient error / declared tolerance:", worst)


def test_mie_coefficient_derivatives_match_arb_and_oracle():
    """da/dx, da/dm, db/dx, db/dm of mie_derivatives.jsonl's coefficient records within the
    declared tolerance of derivatives; every order within twice that tolerance of the oracle's
    closed forms (in the largest derivative's measure)."""
    from scatrs_amd.mie import coefficients

    recs = [r for r in load_golden("mie_derivatives.jsonl") if r["case"].endswith(",coefficients")]
    x = np.array([r["inputs"]["x"] for r in recs])
    m = np.array([_m(r) for r in recs])
    c = coefficients(x, m, derivatives=True)
    stems = (("da_dx", c.d("x").a), ("da_dm", c.d("m").a), ("db_dx", c.d("x").b),
             ("db_dm", c.d("m").b))
    # d/d(Im m) = i d/dm
    assert np.array_equal(c.d("m_imag").a, 1j * c.d("m").a)
    for i, r in enumerate(recs):
        orders = np.array(r["inputs"]["orders"])
        s = oracle.mie_solve(x[i], m[i], derivatives=True)
        for stem, got in stems:
            ref = cvals(r, stem)
            scale = np.max(np.abs(ref))
            if scale == 0.0:
                continue
            err = np.max(np.abs(got[i][orders - 1] - ref)) / scale
            tol = tolerance("coefficient_derivatives", x[i], m[i], 0.0)
            assert err <= max(tol, 1e-9), (r["case"], stem, err, tol)
            full = s[stem]
            n = full.size
            rel = np.max(np.abs(got[i][:n] - full)) / np.max(np.abs(full))
            assert rel <= 2.0 * max(tol, 1e-12), (r["case"], stem, rel, tol)


def test_mie_coefficients_give_the_efficiencies():
    """Q_sca = (2/x^2) sum (2n+1)(|a|^2 + |b|^2) and g of the coefficients equal the
    efficiency entry's on seeded spheres (the same solve, summed on the host here)."""
    from scatrs_amd import workloads
    from scatrs_amd.mie import coefficients, efficiencies

    x, m = workloads.spheres(3000, seed=31)
    x = x[x < 300]
    m = m[: x.size]
    c = coefficients(x, m)
    e = efficiencies(x, m)
    n = np.arange(1, c.a.shape[1] + 1)
    w = 2 * n + 1
    sca = (w * (np.abs(c.a) ** 2 + np.abs(c.b) ** 2)).sum(1)
    q_sca = 2.0 / x ** 2 * sca
    assert np.max(np.abs(q_sca - e.q_sca) / e.q_sca) < 1e-12


@pytest.mark.parametrize("derivatives", [False, True])
def test_layered_coefficients_match_arb(derivatives):
    """Every coefficient record of layered.jsonl inside the validated domain, batched per layer
    count, within the reference's layered tolerance model (layered/domain.rs:96-160)."""
    from scatrs_amd.layered import MAX_LAYERS, coefficients

    recs = layered_records(",coefficients", max_layers=MAX_LAYERS)
    by_l: dict[int, list] = {}
    for r in recs:
        by_l.setdefault(len(r["inputs"]["x"]), []).append(r)
    worst = 0.0
    for L, group in sorted(by_l.items()):
        print(f"layered coefficients: {len(group)} records of {L} layers", flush=True)
        x = np.array([rec_inputs(r)[0] for r in group])
        m = np.array([rec_inputs(r)[1] for r in group])
        n = max(coefficient_n(r) for r in group)
        c = coefficients(x, m, n_max=n, derivatives=derivatives,
                         allow_unvalidated=True)
        assert np.all(c.n_terms == n)
        for i, r in enumerate(group):
            worst = max(worst, check_coefficients(r, c.a[i], c.b[i]))
    print("worst GPU layered coefficient error / tolerance:", worst)


def test_layered_efficiencies_match_arb():
    """Every series record of layered.jsonl inside the validated domain, batched per layer
    count, within the reference's layered tolerance model; with the Jacobian requested the
    values agree to 1e-13 of each output's measure."""
    from scatrs_amd.layered import MAX_LAYERS, efficiencies

    by_l: dict[int, list] = {}
    for r in layered_records(",series", max_layers=MAX_LAYERS):
        by_l.setdefault(len(r["inputs"]["x"]), []).append(r)
    worst = np.zeros(5)
    for L, group in sorted(by_l.items()):
        print(f"layered efficiencies: {len(group)} records of {L} layers", flush=True)
        x = np.array([rec_inputs(r)[0] for r in group])
        m = np.array([rec_inputs(r)[1] for r in group])
        e = efficiencies(x, m)
        ej = efficiencies(x, m, derivatives=True, allow_unvalidated=True)
        q = np.stack([e.q_ext, e.q_sca, e.q_abs, e.q_back, e.g])
        qj = np.stack([ej.q_ext, ej.q_sca, ej.q_abs, ej.q_back, ej.g])
        for i, r in enumerate(group):
            sizes = np.array(efficiency_sizes(q[:, i], x[i, -1]))
            assert np.max(np.abs(q[:, i] - qj[:, i]) / sizes) <= 1e-13, (r["case"], q[:, i],
                                                                        qj[:, i])
            worst = np.maximum(worst, check_efficiencies(r, q[:, i]))
    print("worst GPU layered efficiency error / tolerance:", worst)


def test_layered_derivatives_match_arb():
    """Every layer's da/dx_l, da/dm_l, db/dx_l, db/dm_l of layered_derivatives.jsonl within 1e-9
    of the largest derivative of its kind over the layers (the oracle's own pin)."""
    from scatrs_amd.layered import MAX_LAYERS, coefficients

    for r in load_golden("layered_derivatives.jsonl"):
        x, m = rec_inputs(r)
        if x.size > MAX_LAYERS:
            continue
        orders = np.array(r["inputs"]["orders"])
        c = coefficients(x[None], m[None], n_max=coefficient_n(r), derivatives=True,
                         allow_unvalidated=True)
        for j, stem in enumerate(("da_dx", "da_dm", "db_dx", "db_dm")):
            refs = {l: cvals(r, f"{stem}{l}") for l in range(x.size)
                    if f"{stem}{l}_re" in r["outputs"]}
            scale = max(np.max(np.abs(v)) for v in refs.values())
            for l, ref in refs.items():
                got = c._derivatives[l, j, 0, orders - 1]
                err = np.max(np.abs(got - ref)) / scale
                assert err < 1e-9, (r["case"], stem, l, err)


def test_layered_jacobian_matches_oracle_and_central_differences():
    """Seeded 2-4 layer spheres: the efficiency Jacobian of every layer against the oracle's
    (each layer's coefficient derivatives through the Jacobian sums) to 1e-10 of the largest
    entry of its output, and against central differences of the GPU values (steps 1e-6 x_l and
    1e-6 |m_l|, as the reference's derivative measures step x_l and |m_l|)."""
    from scatrs_amd.layered import efficiencies

    rng = np.random.default_rng(17)
    for L in (2, 3, 4):
        P = 6
        xo = np.exp(rng.uniform(np.log(0.3), np.log(60.0), P))
        fr = np.cumsum(rng.dirichlet(np.full(L, 2.0), P), axis=1)
        x = fr * xo[:, None]
        x[:, -1] = xo
        m = rng.uniform(1.3, 2.0, (P, L)) + 1j * np.exp(rng.uniform(np.log(1e-5), 0.0, (P, L)))
        e = efficiencies(x, m, derivatives=True)
        J = e._jacobian  # (L, 3, 5, P)
        for i in range(P):
            ref = layered_jacobian(x[i], m[i])  # (L, 3, 5)
            got = J[:, :, :, i]
            scale = np.abs(ref).max(axis=(0, 1))
            assert np.max(np.abs(got - ref) / scale) < 1e-10, (L, i)
        # central differences in each layer's x, Re m, Im m of the first particle
        h = 1e-6
        for l in range(L):
            for p in range(3):
                xp, xm, mp, mm = x[:1].copy(), x[:1].copy(), m[:1].copy(), m[:1].copy()
                if p == 0:
                    d = h * x[0, l]
                    xp[0, l] += d
                    xm[0, l] -= d
                elif p == 1:
                    d = h * abs(m[0, l])
                    mp[0, l] += d
                    mm[0, l] -= d
                else:
                    d = min(h * abs(m[0, l]), 0.5 * m[0, l].imag)
                    mp[0, l] += 1j * d
                    mm[0, l] -= 1j * d
                ep, em = efficiencies(xp, mp), efficiencies(xm, mm)
                fd = (np.array([ep.q_ext, ep.q_sca, ep.q_abs, ep.q_back, ep.g]) -
                      np.array([em.q_ext, em.q_sca, em.q_abs, em.q_back, em.g]))[:, 0] / (2 * d)
                got = J[l, p, :, 0]
                scale = np.abs(J[:, :, :, 0]).max(axis=(0, 1))
                assert np.max(np.abs(got - fd) / scale) < 2e-6, (L, l, p, got, fd)


def test_layered_one_layer_is_the_sphere_and_errors():
    """One layer through the layered entry points is the sphere bitwise; the reference's
    messages for bad input (layered/mod.rs:229-319)."""
    from scatrs_amd import InvalidInputError, OutsideValidatedDomainError
    from scatrs_amd.layered import coefficients, efficiencies
    from scatrs_amd.mie import coefficients as mie_coefficients
    from scatrs_amd.mie import efficiencies as mie_efficiencies

    x = np.array([0.3, 12.0, 250.0])
    m = np.array([1.5 + 0.01j, 1.33 + 1e-8j, 1.7 + 0.1j])
    le = efficiencies(x[:, None], m[:, None], derivatives=True)
    se = mie_efficiencies(x, m, derivatives=True)
    for f in ("q_ext", "q_sca", "q_abs", "q_back", "g"):
        assert np.array_equal(getattr(le, f), getattr(se, f))
        assert np.array_equal(getattr(le.d("layer0.x"), f), getattr(se.d("x"), f))
    lc = coefficients(x[:, None], m[:, None])
    sc = mie_coefficients(x, m)
    assert np.array_equal(lc.a, sc.a) and np.array_equal(lc.b, sc.b)
    with pytest.raises(InvalidInputError, match="element 1: the size parameters must be "
                                                "nondecreasing"):
        efficiencies([[1.0, 2.0], [3.0, 2.0]], [[1.5, 1.4], [1.5, 1.4]])
    with pytest.raises(InvalidInputError, match="refractive index of layer 1 must have"):
        efficiencies([[1.0, 2.0]], [[1.5, 1.4 - 0.1j]])
    with pytest.raises(OutsideValidatedDomainError, match="derivatives are validated up to"):
        efficiencies([[500.0, 2000.0]], [[1.5, 1.4]], derivatives=True)
    with pytest.raises(OutsideValidatedDomainError, match="outside the validated domain"):
        efficiencies([[5e3, 2e4]], [[1.5, 1.4]])
    # inside the domain with allow_unvalidated, and finite
    e = efficiencies([[500.0, 2000.0]], [[1.5, 1.4]], derivatives=True, allow_unvalidated=True)
    assert np.all(np.isfinite(e._jacobian))
