"""Checks of layered-sphere results against the reference's Arb records (layered.jsonl) under
its declared tolerance model, shared by the oracle's and the GPU library's tests: the measure
and condition number of each output follow scatrs/tests/common/layered.rs:132-215, with the
oracle's FP64 closed-form derivatives standing in for the reference's double-double ones (they
only set the condition number)."""

from __future__ import annotations

import numpy as np

import oracle
from tests.conftest import load_golden
from tests.tolerance import efficiency_sizes, layered_in_domain, layered_tolerance

Q_KINDS = ("q_ext", "q_sca", "q_abs", "q_back", "g")


def cvals(rec: dict, stem: str) -> np.ndarray:
    return np.array(rec["outputs"][stem + "_re"]) + 1j * np.array(rec["outputs"][stem + "_im"])


def rec_inputs(rec: dict):
    i = rec["inputs"]
    return np.array(i["x"], np.float64), np.array(i["m_re"]) + 1j * np.array(i["m_im"])


def layered_records(kind: str, inside: bool = True, max_layers: int | None = None) -> list[dict]:
    """The layered.jsonl records of one kind (",coefficients" or ",series"), inside the
    reference's validated domain (it refuses the others) when `inside`, of at most max_layers
    layers when given."""
    out = []
    for r in load_golden("layered.jsonl"):
        if not r["case"].endswith(kind):
            continue
        x, m = rec_inputs(r)
        if inside and not layered_in_domain(x, m):
            continue
        if max_layers is not None and x.size > max_layers:
            continue
        out.append(r)
    return out


def coefficient_n(rec: dict) -> int:
    """Enough terms for the record's largest stored order (layered_reference.rs:48-51)."""
    x, _ = rec_inputs(rec)
    return max(oracle.default_n_max(x[-1]), int(max(rec["inputs"]["orders"])))


def check_coefficients(rec: dict, a: np.ndarray, b: np.ndarray, slack: float = 1.0) -> float:
    """Asserts a_n, b_n (index n - 1) at the record's orders within the declared tolerance;
    returns the worst error / tolerance."""
    x, m = rec_inputs(rec)
    orders = np.array(rec["inputs"]["orders"])
    d = oracle.layered_solve(x, m, n=coefficient_n(rec), derivatives=True)["d"]
    worst = 0.0
    for k, (stem, got) in enumerate((("a", a), ("b", b))):
        ref = cvals(rec, stem)
        scale = np.max(np.abs(ref))
        err = np.max(np.abs(got[orders - 1] - ref))
        # sum over layers of x_l |dy/dx_l| + |m_l| |dy/dm_l|, worst over the orders
        num = np.max(sum(x[l] * np.abs(d[l, 2 * k, orders - 1]) +
                         abs(m[l]) * np.abs(d[l, 2 * k + 1, orders - 1]) for l in range(x.size)))
        tol = layered_tolerance("coefficients", x[-1], m, num / scale) * slack
        assert err / scale <= tol, (rec["case"], stem, err / scale, tol)
        worst = max(worst, err / scale / tol)
    return worst


def layered_jacobian(x, m) -> np.ndarray:
    """The oracle's efficiency Jacobian (L, 3, 5) of a layered sphere at the default truncation:
    each layer's group of coefficient derivatives through the sphere's Jacobian sums, with the
    1/x_L^2 normalization's term on the outer layer only."""
    s = oracle.layered_solve(x, m, derivatives=True)
    out = np.zeros((x.size, 3, 5))
    for l in range(x.size):
        g = {"n": s["n"], "a": s["a"], "b": s["b"], "abs_sum": s["abs_sum"],
             "da_dx": s["d"][l, 0].copy(), "da_dm": s["d"][l, 1].copy(),
             "db_dx": s["d"][l, 2].copy(), "db_dm": s["d"][l, 3].copy(),
             "abs_jac": s["abs_jac"][l].copy()}
        out[l] = oracle.efficiencies_jacobian(x[-1], g, outer=(l == x.size - 1))
    return out


def check_efficiencies(rec: dict, q: np.ndarray, slack: float = 1.0) -> np.ndarray:
    """Asserts [Q_ext, Q_sca, Q_abs, Q_back, g] against a series record within the declared
    tolerance; returns the worst error / tolerance of each."""
    x, m = rec_inputs(rec)
    jac = layered_jacobian(x, m)
    sizes = efficiency_sizes(q, x[-1])
    worst = np.zeros(5)
    for k, name in enumerate(Q_KINDS):
        if name not in rec["outputs"]:
            continue
        if k == 2 and np.all(m.imag == 0.0):
            assert q[2] == 0.0, rec["case"]
            continue
        ref = rec["outputs"][name][0]
        num = sum(x[l] * abs(jac[l, 0, k]) + abs(m[l].real) * abs(jac[l, 1, k]) +
                  abs(m[l].imag) * abs(jac[l, 2, k]) for l in range(x.size))
        tol = layered_tolerance(name, x[-1], m, num / sizes[k]) * slack
        err = abs(q[k] - ref) / sizes[k]
        assert err <= tol, (rec["case"], name, q[k], ref, err, tol)
        worst[k] = err / tol
    return worst
