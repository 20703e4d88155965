"""Particles of the path: homogeneous and layered spheres (python/scatrs/particles.py:48-172).

A refractive index is a complex number, an array of one value per wavelength, or a callable of
the wavelength (the reference's refractive-index models are evaluated on the host before the
engine, python/scatrs/particles.py:67-72; any callable returning m(lambda) stands in here).
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Sequence

import numpy as np

from .errors import InvalidInputError

Index = complex | np.ndarray | Callable[[np.ndarray], np.ndarray]


def _at(m: Index, wavelength: np.ndarray) -> np.ndarray:
    if callable(m):
        v = np.asarray(m(wavelength), np.complex128)
    else:
        v = np.asarray(m, np.complex128)
    try:
        return np.broadcast_to(v, wavelength.shape).astype(np.complex128)
    except ValueError:
        raise InvalidInputError(
            f"a refractive index of shape {v.shape} for wavelengths of shape "
            f"{wavelength.shape}") from None


@dataclass(frozen=True)
class Sphere:
    """Homogeneous spheres of refractive index m(lambda)."""

    refractive_index: Index

    layers = 1

    def indices(self, wavelength: np.ndarray) -> np.ndarray:
        """(1,) + wavelength.shape complex."""
        return _at(self.refractive_index, wavelength)[None]

    def fractions(self) -> np.ndarray:
        return np.zeros(0)

    def parameter_names(self) -> tuple[str, ...]:
        return ("m_real", "m_imag")


@dataclass(frozen=True)
class LayeredSphere:
    """A core and concentric shells: layer l's outer radius is radius_fractions[l] times the
    particle's (l < L - 1), its index refractive_index[l] (scatrs/src/ensemble/mod.rs:56-66)."""

    radius_fractions: Sequence[float]
    refractive_index: Sequence[Index]

    def __post_init__(self) -> None:
        fr = np.asarray(self.radius_fractions, np.float64).reshape(-1)
        if fr.size + 1 != len(self.refractive_index):
            raise InvalidInputError(
                f"{len(self.refractive_index)} layers need {len(self.refractive_index) - 1} "
                f"radius fractions, got {fr.size}")
        if np.any(~(fr > 0) | ~(fr <= 1)) or np.any(np.diff(fr) < 0):
            raise InvalidInputError("the radius fractions must be nondecreasing in (0, 1]")

    @property
    def layers(self) -> int:
        return len(self.refractive_index)

    @classmethod
    def coated(cls, core: Index, shell: Index, *, core_fraction: float) -> "LayeredSphere":
        return cls([core_fraction], [core, shell])

    def indices(self, wavelength: np.ndarray) -> np.ndarray:
        """(L,) + wavelength.shape complex."""
        return np.stack([_at(m, wavelength) for m in self.refractive_index])

    def fractions(self) -> np.ndarray:
        return np.asarray(self.radius_fractions, np.float64).reshape(-1)

    def parameter_names(self) -> tuple[str, ...]:
        names = []
        for l in range(self.layers):
            names += [f"layer{l}.m_real", f"layer{l}.m_imag"]
        names += [f"layer{l}.radius_fraction" for l in range(self.layers - 1)]
        return tuple(names)
