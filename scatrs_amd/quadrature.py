"""Gauss-Legendre cosines of the phase-matrix outputs (host, for labelling results)."""

from __future__ import annotations

import numpy as np


def gauss_legendre(n: int):
    """The n-point rule on [-1, 1], nodes ascending."""
    x, w = np.polynomial.legendre.leggauss(int(n))
    return x, w
