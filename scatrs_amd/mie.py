"""Batched homogeneous spheres on the GPU: the efficiencies and their Jacobian, and the Mie
coefficients a_n, b_n and their derivatives.

Mirrors ``scatrs.mie.efficiencies`` and ``scatrs.mie.coefficients`` (python/scatrs/mie.py:
390-446) and the wire layouts of ``_scatrs.mie_efficiencies`` and ``_scatrs.mie_coefficients``
(scatrs-py/src/mie.rs:303-442): values (5, len) in the order Q_ext, Q_sca, Q_abs, Q_back, g and
the Jacobian (3, 5, len) with respect to x, Re m, Im m; coefficients (len, N) zero-padded to
the largest N, derivatives (4, len, N) [da/dx, da/dm, db/dx, db/dm].
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _native

__all__ = ["MieCoefficients", "MieEfficiencies", "coefficients", "coefficients_device",
           "default_n_max", "efficiencies", "efficiencies_device", "efficiencies_host"]

_PARAMS = ("x", "m_real", "m_imag")


def default_n_max(x) -> np.ndarray:
    """N = ceil(x + 6 x^(1/3) + 3) (scatrs/src/mie/mod.rs:122-124)."""
    xs = np.asarray(x, dtype=np.float64)
    n = np.ceil(xs + 6.0 * np.cbrt(xs) + 3.0)
    return n.astype(np.int64)


@dataclass(frozen=True, eq=False)
class MieEfficiencies:
    """Efficiencies and asymmetry parameter, each an array of the broadcast shape of x and m."""

    q_ext: np.ndarray
    q_sca: np.ndarray
    q_abs: np.ndarray
    q_back: np.ndarray
    g: np.ndarray
    params: tuple[str, ...] = ()
    _jacobian: np.ndarray | None = field(default=None, repr=False)

    def d(self, param: str) -> "MieEfficiencies":
        """The derivatives with respect to ``param`` ("x", "m_real" or "m_imag")."""
        if param not in self.params:
            raise ValueError(f"no derivative with respect to {param!r}; computed: {self.params}")
        rows = self._jacobian[_PARAMS.index(param)]
        return MieEfficiencies(*rows)


@dataclass(frozen=True, eq=False)
class MieCoefficients:
    """a_n, b_n of spheres, arrays of shape S + (N,) with N the largest number of terms
    (``a[..., n - 1]`` is a_n, zero beyond each sphere's ``n_terms``); ``d`` gives their
    derivatives (python/scatrs/mie.py:173-212)."""

    a: np.ndarray
    b: np.ndarray
    n_terms: np.ndarray
    params: tuple[str, ...] = ()
    _derivatives: np.ndarray | None = field(default=None, repr=False)

    def d(self, param: str) -> "MieCoefficients":
        """The derivatives with respect to ``param``: "x", "m_real", "m_imag", or "m" for the
        complex derivative d/dm (equal to d/d(Re m); d/d(Im m) is i d/dm)."""
        if param not in self.params and not (param == "m" and self.params):
            raise ValueError(f"no derivative with respect to {param!r}; computed: {self.params}")
        da_dx, da_dm, db_dx, db_dm = self._derivatives
        if param == "x":
            return MieCoefficients(da_dx, db_dx, self.n_terms)
        if param == "m_imag":
            return MieCoefficients(1j * da_dm, 1j * db_dm, self.n_terms)
        return MieCoefficients(da_dm, db_dm, self.n_terms)


def _parameters(derivatives: bool | Sequence[str]) -> tuple[str, ...]:
    if isinstance(derivatives, (bool, np.bool_)):
        return _PARAMS if derivatives else ()
    params = tuple(derivatives)
    for p in params:
        if p not in _PARAMS:
            raise ValueError(f"unknown parameter {p!r}; expected one of {_PARAMS}")
    return params


def efficiencies_device(x, m, *, n_max: int | None = None, derivatives: bool = False,
                        allow_unvalidated: bool = False):
    """Device-resident form: ``x`` a float64 CUDA tensor (len,), ``m`` a complex128 CUDA tensor
    (len,); returns (values (5, len), jacobian (3, 5, len) or None, total terms)."""
    import torch

    if x.dtype != torch.float64 or m.dtype != torch.complex128:
        raise TypeError("x must be float64 and m complex128")
    if x.device.type != "cuda" or m.device != x.device:
        raise ValueError("x and m must be on the same GPU")
    x = x.contiguous().reshape(-1)
    m = m.contiguous().reshape(-1)
    if m.numel() != x.numel():
        raise ValueError("x and m must have the same number of elements")
    n = x.numel()
    ctx = _native.context(x.device.index)
    values = torch.empty((5, n), dtype=torch.float64, device=x.device)
    jac = torch.empty((3, 5, n), dtype=torch.float64, device=x.device) if derivatives else None
    terms = C.c_int64(0)
    rc = ctx.lib.scatrs_mie_efficiencies(
        ctx.handle, C.c_void_p(x.data_ptr()), C.c_void_p(m.data_ptr()), n,
        0 if n_max is None else int(n_max), int(bool(allow_unvalidated)),
        C.c_void_p(values.data_ptr()), C.c_void_p(jac.data_ptr()) if jac is not None else None,
        C.byref(terms))
    ctx.check(rc)
    return values, jac, terms.value


def efficiencies_host(x, m, *, n_max: int | None = None, derivatives: bool = False,
                      allow_unvalidated: bool = False):
    """Host-array form (scatrs_mie_efficiencies_host): ``x`` float64 (len,), ``m`` complex128
    (len,) numpy arrays; returns (values (5, len), jacobian (3, 5, len) or None, total terms)
    as numpy arrays, with one host synchronization and no validation launches."""
    import torch

    x = np.ascontiguousarray(x, np.float64).reshape(-1)
    m = np.ascontiguousarray(m, np.complex128).reshape(-1)
    if m.size != x.size:
        raise ValueError("x and m must have the same number of elements")
    n = x.size
    ctx = _native.context(torch.cuda.current_device())
    values = np.empty((5, n))
    jac = np.empty((3, 5, n)) if derivatives else None
    terms = C.c_int64(0)
    rc = ctx.lib.scatrs_mie_efficiencies_host(
        ctx.handle, C.c_void_p(x.ctypes.data), C.c_void_p(m.ctypes.data), n,
        0 if n_max is None else int(n_max), int(bool(allow_unvalidated)),
        C.c_void_p(values.ctypes.data), C.c_void_p(jac.ctypes.data) if jac is not None else None,
        C.byref(terms))
    ctx.check(rc)
    return values, jac, terms.value


def efficiencies(x, m, *, n_max: int | None = None, derivatives: bool | Sequence[str] = False,
                 allow_unvalidated: bool = False) -> MieEfficiencies:
    """Efficiencies and asymmetry parameter of spheres (host arrays in, host arrays out)."""
    params = _parameters(derivatives)
    xb, mb = np.broadcast_arrays(np.asarray(x, dtype=np.float64),
                                 np.asarray(m, dtype=np.complex128))
    shape = xb.shape
    v, jac, _ = efficiencies_host(xb, mb, n_max=n_max, derivatives=bool(params),
                                  allow_unvalidated=allow_unvalidated)
    fields = [v[k].reshape(shape) for k in range(5)]
    j = None if jac is None else jac.reshape((3, 5, *shape))
    return MieEfficiencies(*fields, params=params, _jacobian=j)


def _device_pair(x, m):
    import torch

    if x.dtype != torch.float64 or m.dtype != torch.complex128:
        raise TypeError("x must be float64 and m complex128")
    if x.device.type != "cuda" or m.device != x.device:
        raise ValueError("x and m must be on the same GPU")
    return x.contiguous(), m.contiguous()


def coefficients_device(x, m, *, n_max: int | None = None, derivatives: bool = False,
                        allow_unvalidated: bool = False):
    """Device-resident form: ``x`` float64 (len,), ``m`` complex128 (len,) CUDA tensors; returns
    (n_terms int64 (len,), a, b complex128 (len, N), derivatives complex128 (4, len, N) or
    None), N the largest number of terms (scatrs_mie_terms, then scatrs_mie_coefficients)."""
    import torch

    x, m = _device_pair(x, m)
    x, m = x.reshape(-1), m.reshape(-1)
    if m.numel() != x.numel():
        raise ValueError("x and m must have the same number of elements")
    n = x.numel()
    ctx = _native.context(x.device.index)
    nm = 0 if n_max is None else int(n_max)
    width, total = C.c_int64(0), C.c_int64(0)
    ctx.check(ctx.lib.scatrs_mie_terms(ctx.handle, C.c_void_p(x.data_ptr()),
                                       C.c_void_p(m.data_ptr()), n, nm, int(bool(allow_unvalidated)),
                                       C.byref(width), C.byref(total)))
    w = max(width.value, 1)
    dev = x.device
    terms = torch.empty(n, dtype=torch.int64, device=dev)
    a = torch.empty((n, w), dtype=torch.complex128, device=dev)
    b = torch.empty((n, w), dtype=torch.complex128, device=dev)
    d = torch.empty((4, n, w), dtype=torch.complex128, device=dev) if derivatives else None
    ctx.check(ctx.lib.scatrs_mie_coefficients(
        ctx.handle, C.c_void_p(x.data_ptr()), C.c_void_p(m.data_ptr()), n, nm,
        int(bool(allow_unvalidated)), w, C.c_void_p(terms.data_ptr()), C.c_void_p(a.data_ptr()),
        C.c_void_p(b.data_ptr()), C.c_void_p(d.data_ptr()) if d is not None else None,
        C.byref(total)))
    return terms, a, b, d


def coefficients(x, m, *, n_max: int | None = None, derivatives: bool | Sequence[str] = False,
                 allow_unvalidated: bool = False) -> MieCoefficients:
    """Mie coefficients a_n and b_n of spheres (host arrays in, host arrays out), shape
    S + (N,) zero-padded to the largest number of terms N (python/scatrs/mie.py:418-446)."""
    import torch

    params = _parameters(derivatives)
    xb, mb = np.broadcast_arrays(np.asarray(x, dtype=np.float64),
                                 np.asarray(m, dtype=np.complex128))
    shape = xb.shape
    dev = torch.device("cuda", torch.cuda.current_device())
    xt = torch.from_numpy(np.ascontiguousarray(xb).reshape(-1)).to(dev)
    mt = torch.from_numpy(np.ascontiguousarray(mb).reshape(-1)).to(dev)
    terms, a, b, d = coefficients_device(xt, mt, n_max=n_max, derivatives=bool(params),
                                         allow_unvalidated=allow_unvalidated)
    w = a.shape[1]
    dd = None if d is None else d.cpu().numpy().reshape((4, *shape, w))
    return MieCoefficients(a.cpu().numpy().reshape((*shape, w)),
                           b.cpu().numpy().reshape((*shape, w)),
                           terms.cpu().numpy().reshape(shape), params, dd)
