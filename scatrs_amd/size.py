"""Size distributions of the path (python/scatrs/size.py:103-122): number densities of the
radius, with array-valued parameters forming a grid of distributions."""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .errors import InvalidInputError

LOGNORMAL, GAMMA, MODIFIED_GAMMA, POWER_LAW, MONODISPERSE, SEGMENT = 0, 1, 2, 3, 4, 5
_ZERO = np.float64(0.0)


@dataclass(frozen=True)
class LogNormal:
    """n(r) ∝ exp(-(ln r - ln r_g)^2 / (2 ln^2 sigma_g)) / r: median radius r_g > 0, width
    sigma_g > 1 (scatrs/src/ensemble/distribution.rs:17-23)."""

    median_radius: float | np.ndarray
    width: float | np.ndarray

    kind = LOGNORMAL
    names = ("median_radius", "width")

    def params(self):
        """(p0, p1, p2) as the library takes them."""
        return (np.asarray(self.median_radius, np.float64), np.asarray(self.width, np.float64),
                _ZERO)


@dataclass(frozen=True)
class Gamma:
    """Hansen & Travis's gamma distribution n(r) ∝ r^((1-3b)/b) exp(-r/(a b)): effective radius
    a > 0, effective variance 0 < b < 1/2 (distribution.rs:24-31)."""

    effective_radius: float | np.ndarray
    effective_variance: float | np.ndarray

    kind = GAMMA
    names = ("effective_radius", "effective_variance")

    def params(self):
        """(p0, p1, p2) as the library takes them."""
        return (np.asarray(self.effective_radius, np.float64),
                np.asarray(self.effective_variance, np.float64), _ZERO)


@dataclass(frozen=True)
class ModifiedGamma:
    """Deirmendjian's modified gamma distribution n(r) ∝ r^alpha exp(-b r^gamma): alpha > -1,
    b > 0, gamma > 0 (distribution.rs:32-40)."""

    alpha: float | np.ndarray
    b: float | np.ndarray
    gamma: float | np.ndarray

    kind = MODIFIED_GAMMA
    names = ("alpha", "b", "gamma")

    def params(self):
        """(p0, p1, p2) as the library takes them."""
        return tuple(np.asarray(v, np.float64) for v in (self.alpha, self.b, self.gamma))


@dataclass(frozen=True)
class PowerLaw:
    """n(r) ∝ r^(-exponent) on [min_radius, max_radius], the limits fixed
    (distribution.rs:41-49)."""

    exponent: float | np.ndarray
    min_radius: float | np.ndarray
    max_radius: float | np.ndarray

    kind = POWER_LAW
    names = ("exponent",)

    def params(self):
        """(p0, p1, p2) as the library takes them."""
        return tuple(np.asarray(v, np.float64)
                     for v in (self.exponent, self.min_radius, self.max_radius))


@dataclass(frozen=True)
class Monodisperse:
    """All spheres of one radius > 0 (distribution.rs:50-54)."""

    radius: float | np.ndarray

    kind = MONODISPERSE
    names = ("radius",)

    def params(self):
        """(p0, p1, p2) as the library takes them."""
        return np.asarray(self.radius, np.float64), _ZERO, _ZERO


@dataclass(frozen=True)
class Tabulated:
    """n(r) linear between the points (radius[i], density[i]) and 0 outside them: radii
    increasing and positive, densities >= 0 and not all 0; data, not parameters
    (distribution.rs:55-61,149-170). Integrated segment by segment, each a linear density on
    its interval (the library's segment kind), and mixed by the segments' number integrals
    (d_i + d_i+1)/2 (r_i+1 - r_i), so that the kinks at the points need no quadrature."""

    radius: tuple
    density: tuple

    names = ()

    def __post_init__(self) -> None:
        r = np.asarray(self.radius, np.float64).reshape(-1)
        d = np.asarray(self.density, np.float64).reshape(-1)
        if r.size < 2 or r.size != d.size:
            raise InvalidInputError(
                "a tabulated distribution needs at least 2 points and as many densities as "
                f"radii, got {r.size} radii and {d.size} densities")
        for v in r:
            if not (np.isfinite(v) and v > 0.0):
                raise InvalidInputError(f"radius must be finite and > 0, got {v}")
        if np.any(np.diff(r) <= 0.0):
            raise InvalidInputError("tabulated radii must increase")
        if not np.all(np.isfinite(d) & (d >= 0.0)):
            raise InvalidInputError("tabulated densities must be finite and >= 0")
        if not np.any(d > 0.0):
            raise InvalidInputError("tabulated densities must not all be 0")
        object.__setattr__(self, "radius", tuple(float(v) for v in r))
        object.__setattr__(self, "density", tuple(float(v) for v in d))

    def segments(self):
        """The segments with any density: (r_0, r_1, tilt d_1 / (d_0 + d_1)) and their number
        integrals."""
        out, w = [], []
        for r0, r1, d0, d1 in zip(self.radius, self.radius[1:], self.density, self.density[1:]):
            if d0 + d1 > 0.0:
                out.append((r0, r1, d1 / (d0 + d1)))
                w.append(0.5 * (d0 + d1) * (r1 - r0))
        return out, np.array(w)


_SINGLE = (LogNormal, Gamma, ModifiedGamma, PowerLaw, Monodisperse)


@dataclass(frozen=True)
class MultiModal:
    """sum_j w_j n_j(r) with each mode n_j normalized: ``modes`` of (weight w_j > 0, a
    distribution other than MultiModal with scalar parameters)
    (scatrs/src/ensemble/distribution.rs:62-67,172-180). Each
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
            if not isinstance(d, _SINGLE + (Tabulated,)):
                raise TypeError(f"mode {j} must be a LogNormal, Gamma, ModifiedGamma, PowerLaw, "
                                f"Monodisperse or Tabulated distribution, not "
                                f"{type(d).__name__}")
            if not isinstance(d, Tabulated) and any(np.ndim(p) for p in d.params()):
                raise InvalidInputError(f"mode {j}: a mode takes scalar parameters")
            out.append((w, d))
        object.__setattr__(self, "modes", tuple(out))

    def fractions(self) -> np.ndarray:
        """The modes' number fractions w_j / sum w."""
        w = np.array([w for w, _ in self.modes])
        return w / w.sum()


Distribution = (LogNormal | Gamma | ModifiedGamma | PowerLaw | Monodisperse | Tabulated |
                MultiModal)


def check(d: object, what: str = "size") -> Distribution:
    if isinstance(d, (MultiModal, Tabulated)):
        return d
    if not isinstance(d, _SINGLE):
        raise TypeError(f"{what} must be a LogNormal, Gamma, ModifiedGamma, PowerLaw, "
                        f"Monodisperse, Tabulated or MultiModal distribution, not "
                        f"{type(d).__name__}")
    try:
        np.broadcast(*d.params())
    except ValueError as e:
        raise InvalidInputError(f"{what}: parameters do not broadcast: {e}") from None
    return d
