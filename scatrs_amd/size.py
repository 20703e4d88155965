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


@dataclass(frozen=True)
class MultiModal:
    """sum_j w_j n_j(r) with each mode n_j normalized: ``modes`` of (weight w_j > 0, LogNormal
    or Gamma with scalar parameters) (scatrs/src/ensemble/distribution.rs:62-67,172-180). Each
    mode is integrated on its own and the modes' averages are mixed with their number
    fractions w_j / sum w (scatrs/src/ensemble/engine.rs:143-160,2120-2150)."""

    modes: tuple

    def __post_init__(self) -> None:
        try:
            modes = tuple(self.modes)
        except TypeError:
            raise TypeError("modes must be a sequence of (weight, distribution) pairs") from None
        if not modes:
            raise InvalidInputError("a multimodal distribution needs a mode")
        out = []
        for j, mode in enumerate(modes):
            try:
                w, d = mode
            except (TypeError, ValueError):
                raise TypeError(f"mode {j} must be a (weight, distribution) pair") from None
            w = float(w)
            if not (np.isfinite(w) and w > 0.0):
                raise InvalidInputError(f"the weight of mode {j} must be finite and > 0, got {w}")
            if not isinstance(d, (LogNormal, Gamma)):
                raise TypeError(f"mode {j} must be a LogNormal or a Gamma distribution, not "
                                f"{type(d).__name__}")
            if any(np.ndim(p) for p in d.params()):
                raise InvalidInputError(f"mode {j}: a mode takes scalar parameters")
            out.append((w, d))
        object.__setattr__(self, "modes", tuple(out))

    def fractions(self) -> np.ndarray:
        """The modes' number fractions w_j / sum w."""
        w = np.array([w for w, _ in self.modes])
        return w / w.sum()


Distribution = LogNormal | Gamma | MultiModal


def check(d: object, what: str = "size") -> LogNormal | Gamma | MultiModal:
    if isinstance(d, MultiModal):
        return d
    if not isinstance(d, (LogNormal, Gamma)):
        raise TypeError(f"{what} must be a LogNormal, Gamma or MultiModal distribution, "
                        f"not {type(d).__name__}")
    p0, p1 = d.params()
    try:
        np.broadcast(p0, p1)
    except ValueError as e:
        raise InvalidInputError(f"{what}: parameters do not broadcast: {e}") from None
    return d
