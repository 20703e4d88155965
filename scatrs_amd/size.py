"""Size distributions of the path (python/scatrs/size.py:103-122): number densities of the
radius, with array-valued parameters forming a grid of distributions."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .errors import InvalidInputError

LOGNORMAL, GAMMA = 0, 1


@dataclass(frozen=True)
class LogNormal:
    """n(r) ∝ exp(-(ln r - ln r_g)^2 / (2 ln^2 sigma_g)) / r: median radius r_g > 0, width
    sigma_g > 1 (scatrs/src/ensemble/distribution.rs:17-23)."""

    median_radius: float | np.ndarray
    width: float | np.ndarray

    kind = LOGNORMAL
    names = ("median_radius", "width")

    def params(self):
        return np.asarray(self.median_radius, np.float64), np.asarray(self.width, np.float64)


@dataclass(frozen=True)
class Gamma:
    """Hansen & Travis's gamma distribution n(r) ∝ r^((1-3b)/b) exp(-r/(a b)): effective radius
    a > 0, effective variance 0 < b < 1/2 (distribution.rs:24-31)."""

    effective_radius: float | np.ndarray
    effective_variance: float | np.ndarray

    kind = GAMMA
    names = ("effective_radius", "effective_variance")

    def params(self):
        return (np.asarray(self.effective_radius, np.float64),
                np.asarray(self.effective_variance, np.float64))


Distribution = LogNormal | Gamma


def check(d: object, what: str = "size") -> LogNormal | Gamma:
    if not isinstance(d, (LogNormal, Gamma)):
        raise TypeError(f"{what} must be a LogNormal or a Gamma distribution, "
                        f"not {type(d).__name__}")
    p0, p1 = d.params()
    try:
        np.broadcast(p0, p1)
    except ValueError as e:
        raise InvalidInputError(f"{what}: parameters do not broadcast: {e}") from None
    return d
