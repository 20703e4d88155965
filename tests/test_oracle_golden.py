"""The CPU oracle against the reference's Arb golden vectors, under the reference's own
declared tolerance model (scatrs/tests/mie_reference.rs:52-300; layered_reference.rs;
mie_greek_reference.rs; gauss_legendre.rs; wigner_d.rs). These pin the oracle before it is
trusted as the GPU library's checker."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from tests.conftest import load_golden
from tests.tolerance import EPS, efficiency_conditions, efficiency_sizes, tolerance

NAMES = ["q_ext", "q_sca", "q_abs", "q_back", "g"]


def _cvals(rec: dict, stem: str) -> np.ndarray:
    return np.array(rec["outputs"][stem + "_re"]) + 1j * np.array(rec["outputs"][stem + "_im"])


def _m(rec: dict):
    i = rec["inputs"]
    if isinstance(i["m_re"], list):
        return np.array(i["m_re"]) + 1j * np.array(i["m_im"])
    return complex(i["m_re"], i["m_im"])


def test_coefficients_match_arb():
    worst = 0.0
    for r in load_golden("mie.jsonl"):
        if not r["case"].endswith(",coefficients"):
            continue
        x, m = r["inputs"]["x"], _m(r)
        orders = np.array(r["inputs"]["orders"])
        s = oracle.mie_solve(x, m, derivatives=True)
        for stem, got, dx, dm in (("a", s["a"], s["da_dx"], s["da_dm"]),
                                  ("b", s["b"], s["db_dx"], s["db_dm"])):
            ref = _cvals(r, stem)
            g = got[orders - 1]
            scale = np.max(np.abs(ref))
            err = np.max(np.abs(g - ref)) / scale
            cond = np.max(x * np.abs(dx[orders - 1]) + abs(m) * np.abs(dm[orders - 1])) / scale
            tol = tolerance("coefficients", x, m, cond)
            worst = max(worst, err / tol)
            assert err <= tol, (r["case"], stem, err, tol)
    print("worst coefficient error / tolerance:", worst)


def test_efficiencies_and_amplitudes_match_arb():
    worst = np.zeros(6)
    for r in load_golden("mie.jsonl"):
        if not r["case"].endswith(",series"):
            continue
        x, m = r["inputs"]["x"], _m(r)
        s = oracle.mie_solve(x, m, derivatives=True)
        q = oracle.efficiencies(x, s)
        jac = oracle.efficiencies_jacobian(x, s)
        sizes = efficiency_sizes(q, x)
        conds = efficiency_conditions(q, jac, x, m)
        for k, name in enumerate(NAMES):
            ref = r["outputs"][name][0]
            if k == 2 and m.imag == 0.0:
                assert q[2] == 0.0, r["case"]
                continue
            err = abs(q[k] - ref) / sizes[k]
            tol = tolerance("efficiencies", x, m, conds[k])
            worst[k] = max(worst[k], err / tol)
            assert err <= tol, (r["case"], name, q[k], ref, err, tol)
        mu = np.array(r["inputs"]["mu"])
        s1, s2 = oracle.amplitudes(s, mu)
        dx = {"n": s["n"], "a": s["da_dx"], "b": s["db_dx"]}
        dm = {"n": s["n"], "a": s["da_dm"], "b": s["db_dm"]}
        d1x, d2x = oracle.amplitudes(dx, mu)
        d1m, d2m = oracle.amplitudes(dm, mu)
        s_rms = x * np.sqrt(q[1]) / 2.0
        for stem, got, gx, gm in (("s1", s1, d1x, d1m), ("s2", s2, d2x, d2m)):
            ref = _cvals(r, stem)
            for j in range(mu.size):
                scale = max(abs(ref[j]), s_rms)
                cond = (x * abs(gx[j]) + abs(m) * abs(gm[j])) / scale
                err = abs(got[j] - ref[j]) / scale
                tol = tolerance("amplitudes", x, m, cond)
                worst[5] = max(worst[5], err / tol)
                assert err <= tol, (r["case"], stem, mu[j], err, tol)
    print("worst error / tolerance (Q_ext, Q_sca, Q_abs, Q_back, g, S):", worst)


def test_derivatives_match_arb():
    worst = 0.0
    for r in load_golden("mie_derivatives.jsonl"):
        if not r["case"].endswith(",coefficients"):
            continue
        x, m = r["inputs"]["x"], _m(r)
        orders = np.array(r["inputs"]["orders"])
        s = oracle.mie_solve(x, m, derivatives=True)
        for stem in ("da_dx", "da_dm", "db_dx", "db_dm"):
            ref = _cvals(r, stem)
            got = s[stem][orders - 1]
            scale = np.max(np.abs(ref))
            if scale == 0.0:
                continue
            err = np.max(np.abs(got - ref)) / scale
            # The condition of the derivatives needs second derivatives; the prior term
            # (1 + max(1,|m|) x) of the reference's model carries the bound here, with a
            # relative floor of the same model's constant.
            tol = tolerance("coefficient_derivatives", x, m, 0.0)
            worst = max(worst, err / tol)
            assert err <= max(tol, 1e-9), (r["case"], stem, err, tol)
    print("worst derivative error / tolerance:", worst)


def test_layered_match_arb():
    """a_n, b_n and the efficiencies of every layered.jsonl record inside the validated domain
    (the reference refuses the rest) within the reference's layered tolerance model
    C (kappa' + cond) 2^-52 (layered/domain.rs:96-160, layered_reference.rs:147-265)."""
    from tests.golden_checks import (check_coefficients, check_efficiencies, coefficient_n,
                                     layered_records, rec_inputs)

    worst_c, worst_q = 0.0, np.zeros(5)
    recs = layered_records(",coefficients")
    assert len(recs) > 300
    for r in recs:
        x, m = rec_inputs(r)
        s = oracle.layered_solve(x, m, n=coefficient_n(r))
        worst_c = max(worst_c, check_coefficients(r, s["a"], s["b"]))
    for r in layered_records(",series"):
        x, m = rec_inputs(r)
        s = oracle.layered_solve(x, m)
        worst_q = np.maximum(worst_q, check_efficiencies(r, oracle.efficiencies(x[-1], s)))
    print("worst layered error / tolerance: coefficients", worst_c, "efficiencies", worst_q)


def test_layered_derivatives_match_arb():
    for r in load_golden("layered_derivatives.jsonl"):
        x = np.array(r["inputs"]["x"])
        m = _m(r)
        orders = np.array(r["inputs"]["orders"])
        s = oracle.layered_solve(x, m, derivatives=True)
        for j, stem in enumerate(("da_dx", "da_dm", "db_dx", "db_dm")):
            refs = {l: _cvals(r, f"{stem}{l}") for l in range(x.size)
                    if f"{stem}{l}_re" in r["outputs"]}
            # Measured against the largest derivative of this kind over every layer: a
            # zero-thickness shell's own derivative is exactly 0 to within 1e-155.
            scale = max(np.max(np.abs(v)) for v in refs.values())
            for l, ref in refs.items():
                got = s["d"][l, j, orders - 1]
                err = np.max(np.abs(got - ref)) / scale
                assert err < 1e-9, (r["case"], stem, l, err)


def test_one_layer_is_the_sphere_bitwise():
    for x, m in ((0.3, 1.5 + 0.01j), (12.0, 1.33 + 1e-8j), (250.0, 1.7 + 0.1j)):
        s = oracle.mie_solve(x, m, derivatives=True)
        l = oracle.layered_solve([x], [m], derivatives=True)
        assert np.array_equal(s["a"], l["a"]) and np.array_equal(s["b"], l["b"])
        assert s["abs_sum"] == l["abs_sum"]


def test_equal_indices_give_the_sphere():
    x, m = 7.5, 1.45 + 0.002j
    s = oracle.mie_solve(x, m)
    l = oracle.layered_solve([3.0, 5.0, x], [m, m, m])
    assert np.max(np.abs(l["a"] - s["a"])) <= 1e-13 * np.max(np.abs(s["a"]))
    assert np.max(np.abs(l["b"] - s["b"])) <= 1e-13 * np.max(np.abs(s["b"]))


def test_gauss_legendre_match_arb():
    for r in load_golden("gauss_legendre.jsonl"):
        n = r["inputs"]["n"]
        nodes, weights = oracle.gauss_legendre(n)
        idx = np.array(r["inputs"].get("indices", list(range(n))))
        rn = np.array(r["outputs"]["nodes"])
        rw = np.array(r["outputs"]["weights"])
        assert np.max(np.abs(nodes[idx] - rn)) <= 4 * EPS, n
        assert np.max(np.abs(weights[idx] - rw) / rw) <= 32 * EPS * max(1, np.log2(n)), n


def test_wigner_d_match_arb():
    for r in load_golden("wigner_d.jsonl"):
        i = r["inputs"]
        got = oracle.wigner_d(i["m"], i["n"], i["x"], i["l_max"])
        ref = np.array(r["outputs"]["d"])
        ls = np.array(i["ls"]) if "ls" in i else np.arange(i["l_max"] + 1)
        err = np.abs(got[ls] - ref)
        # The reference's declared f64 bound: 16 (l + 1) 2^-52 absolute (wigner.rs:65-71).
        assert np.all(err <= 16 * (ls + 1) * EPS), (r["case"], np.max(err / ((ls + 1) * EPS)))


def test_sphere_greek_match_arb():
    worst = 0.0
    for r in load_golden("mie_greek.jsonl"):
        i = r["inputs"]
        if "mu" in i:
            continue
        x, m, order = i["x"], complex(i["m_re"], i["m_im"]), i["order"]
        s = oracle.mie_solve(x, m)
        n = s["n"]
        projected = min(order, 2 * n)
        n_q = n + (projected + 2) // 2
        nodes, weights = oracle.gauss_legendre(n_q)
        s1, s2 = oracle.amplitudes(s, nodes)
        c = 4.0 / (x * x)
        f11 = c * 0.5 * (np.abs(s1) ** 2 + np.abs(s2) ** 2)
        f12 = c * 0.5 * (np.abs(s2) ** 2 - np.abs(s1) ** 2)
        p = s2 * np.conj(s1)
        g = oracle.greek_project(nodes, weights, f11, f12, c * p.real, c * p.imag, projected)
        g = g / g[0, 0]
        g[0, 0] = 1.0
        names = ["alpha1", "alpha2", "alpha3", "alpha4", "beta1", "beta2"]
        for k, name in enumerate(names):
            if name not in r["outputs"]:
                continue
            ref = np.array(r["outputs"][name])
            got = np.zeros(ref.size)
            got[: min(ref.size, projected + 1)] = g[k, : min(ref.size, projected + 1)]
            err = np.max(np.abs(got - ref)) / np.max(np.abs(np.array(r["outputs"]["alpha1"])))
            worst = max(worst, err)
            assert err <= 1e-10, (r["case"], name, err)
    print("worst Greek error (of max |alpha1|):", worst)
