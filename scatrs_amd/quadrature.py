"""The library's Gauss-Legendre rules (scatrs_gauss_legendre): the cosines at which the phase
matrix is given for n_angles, as the kernels use them (host only, no device)."""

from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native
from .errors import InvalidInputError

MAX_POINTS = 65536  # the largest rule the library builds


def gauss_legendre(n: int):
    """The n-point rule on [-1, 1] (nodes ascending, exactly symmetric), bit for bit the one
    the kernels use; 1 <= n <= 65536."""
    n = int(n)
    if not 1 <= n <= MAX_POINTS:
        raise InvalidInputError(f"Gauss-Legendre rule needs 1 <= n <= {MAX_POINTS}, got {n}")
    x = np.empty(n)
    w = np.empty(n)
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    if _native.load().scatrs_gauss_legendre(n, ptr(x), ptr(w)) != 0:
        raise InvalidInputError(f"Gauss-Legendre rule of {n} points")
    return x, w
