"""The reference's declared accuracy model (mie/domain.rs:107-199; tests/common/mie.rs:60-200).

Every output is within C * (kappa' + cond) * 2^-52 of the exact value in the measure of its
output: kappa = (1 + max(1, |m|) x) * max(1, |m|/|m - 1|), cond the relative condition number
of the output with respect to x and m, C a per-output constant. The reference computes cond
from double-double derivatives; these helpers take the FP64 closed-form derivatives instead,
which are accurate to many digits, ample for a condition number.
"""

from __future__ import annotations

import numpy as np

EPS = 2.0 ** -52

CONSTANT = {
    "coefficients": 16.0,
    "efficiencies": 16.0,
    "amplitudes": 128.0,
    "amplitude_derivatives": 64.0,
    "jacobian": 256.0,
    "coefficient_derivatives": 512.0,
    "phase": 32.0,
    "greek": 8.0,
}


def near_one(m: complex) -> float:
    return max(abs(m) / abs(m - 1.0), 1.0) if m != 1.0 else float("inf")


def kappa(x: float, m: complex) -> float:
    return (1.0 + max(abs(m), 1.0) * x) * near_one(m)


def tolerance(kind: str, x: float, m: complex, cond: float, order: int | None = None) -> float:
    k = kappa(x, m)
    if kind in ("coefficient_derivatives", "amplitude_derivatives"):
        prior = 1.0 + max(abs(m), 1.0) * x
    elif kind == "jacobian":
        prior = k * near_one(m)
    elif kind == "greek":
        prior = k + ((order + 1) ** 2 + x * x) / 8.0
    else:
        prior = k
    return CONSTANT[kind] * (prior + cond) * EPS


def g_floor(x: float) -> float:
    return 1e-3 * min(x * x, 1.0)


def efficiency_sizes(q: np.ndarray, x: float) -> list[float]:
    """Denominators of the Q_ext, Q_sca, Q_abs, Q_back, g measures (tests/common/mie.rs:137)."""
    q_abs = q[2] if q[2] > 0 else q[0]
    return [q[0], q[1], q_abs, max(q[3], q[1]), max(abs(q[4]), g_floor(x))]


def efficiency_conditions(q: np.ndarray, jac: np.ndarray, x: float, m: complex) -> list[float]:
    sizes = efficiency_sizes(q, x)
    return [
        (x * abs(jac[0, k]) + abs(m.real) * abs(jac[1, k]) + abs(m.imag) * abs(jac[2, k]))
        / sizes[k]
        for k in range(5)
    ]


# ---- layered spheres (scatrs/src/layered/domain.rs:96-160; tests/common/layered.rs:132-215) ----

LAYERED_CONSTANT = {
    "coefficients": 32.0,
    "q_ext": 32.0,
    "q_sca": 16.0,
    "q_abs": 256.0,
    "q_back": 64.0,
    "g": 16.0,
    "amplitudes": 128.0,
    "coefficient_derivatives": 1024.0,
    "jacobian": 256.0,
}


def layered_kappa(x_outer: float, m) -> float:
    """(1 + max_l max(1, |m_l|) x_L) max_l max(1, |m_l|/|m_l - 1|) over the layers whose index
    is not exactly 1 (infinite when every layer has m = 1)."""
    ms = np.atleast_1d(np.asarray(m, np.complex128))
    big = max(1.0, float(np.max(np.abs(ms))))
    ones = [max(abs(v) / abs(v - 1.0), 1.0) for v in ms if v != 1.0]
    near = max(ones) if ones else float("inf")
    return (1.0 + big * x_outer) * near


def layered_tolerance(kind: str, x_outer: float, m, cond: float) -> float:
    """C (kappa' + cond) 2^-52 with the layered kappa (layered/domain.rs:115-132)."""
    k = layered_kappa(x_outer, m)
    ms = np.atleast_1d(np.asarray(m, np.complex128))
    big = max(1.0, float(np.max(np.abs(ms))))
    near = max(k / (1.0 + big * x_outer), 1.0)
    prior = k * near if kind == "jacobian" else k
    return LAYERED_CONSTANT[kind] * (prior + cond) * EPS


def layered_in_domain(x, m, derivatives: bool = False) -> bool:
    """The reference's validated domain of layered spheres (layered/domain.rs:53-95); one
    layer is a sphere, with the sphere's domain (mie/domain.rs:123-140)."""
    x = np.atleast_1d(np.asarray(x, np.float64))
    ms = np.atleast_1d(np.asarray(m, np.complex128))
    if x.size == 1:
        v = ms[0]
        return (1e-6 <= x[0] <= 2e4 and 0.0 < v.real <= 10.0 and 0.0 <= v.imag <= 10.0
                and abs(v) >= 0.1)
    outer = x[-1]
    ok = (x.size <= 16 and 1e-6 <= outer <= 1e4
          and all(abs(v) * outer <= 3e4 and 0.0 < v.real <= 10.0 and 0.0 <= v.imag <= 10.0
                  and abs(v) >= 0.1 for v in ms))
    return ok and (not derivatives or outer <= 1e3)
