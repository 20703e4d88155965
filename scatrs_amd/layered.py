"""Batched layered spheres on the GPU: the efficiencies with every layer's Jacobian, and the Mie
coefficients a_n, b_n with every layer's derivatives.

Mirrors ``scatrs.layered.efficiencies`` and ``scatrs.layered.coefficients``
(python/scatrs/layered.py:240-296) and the wire layouts of ``_scatrs.layered_efficiencies`` and
``_scatrs.layered_coefficients`` (scatrs-py/src/layered.rs:123-267): x and m of shape S + (L,),
one value per layer, core first; values (5, P), Jacobian (L, 3, 5, P) with respect to each
layer's x_l, Re m_l, Im m_l; coefficients (P, N) zero-padded to the largest N, derivatives
(L, 4, P, N) [da/dx_l, da/dm_l, db/dx_l, db/dm_l].
"""

from __future__ import annotations

import ctypes as C
import re
from dataclasses import dataclass
from typing import Sequence

import numpy as np

from . import _native
from .mie import MieCoefficients, MieEfficiencies
from .mie import default_n_max as _sphere_default_n_max

__all__ = ["LayeredCoefficients", "LayeredEfficiencies", "MAX_LAYERS", "coefficients",
           "coefficients_device", "default_n_max", "efficiencies", "efficiencies_device",
           "parameters"]

# Layers per particle this library solves (the reference accepts up to 2^10; its validated
# domain ends at 16).
MAX_LAYERS = 8

_PARTS = ("x", "m_real", "m_imag")
_NAME = re.compile(r"layer(\d+)\.(x|m_real|m_imag|m)")


def parameters(layers: int) -> tuple[str, ...]:
    """The parameter names of a particle of ``layers`` layers, in the order of the results
    (python/scatrs/layered.py:67-69)."""
    return tuple(f"layer{j}.{p}" for j in range(layers) for p in _PARTS)


def default_n_max(x) -> np.ndarray:
    """The sphere's default number of terms at the outer size parameter ``x[..., -1]``."""
    xs = np.asarray(x, dtype=np.float64)
    if xs.ndim == 0 or xs.shape[-1] == 0:
        raise ValueError("x needs a trailing axis of at least one layer")
    return _sphere_default_n_max(xs[..., -1])


def _inputs(x, m):
    """x and m broadcast to one shape S + (L,), as C-contiguous (P, L) arrays."""
    xs = np.asarray(x, dtype=np.float64)
    ms = np.asarray(m, dtype=np.complex128)
    if xs.ndim == 0 or ms.ndim == 0:
        raise ValueError("x and m need a trailing axis of layers (one value per layer)")
    if xs.shape[-1] != ms.shape[-1]:
        raise ValueError(f"x and m must have one value per layer: their last axes have "
                         f"{xs.shape[-1]} and {ms.shape[-1]}")
    if xs.shape[-1] == 0:
        raise ValueError("a layered sphere needs at least one layer")
    xb, mb = np.broadcast_arrays(xs, ms)
    layers = xb.shape[-1]
    return (np.ascontiguousarray(xb).reshape(-1, layers),
            np.ascontiguousarray(mb).reshape(-1, layers), xb.shape[:-1])


def _parameters(derivatives: bool | Sequence[str], layers: int) -> tuple[str, ...]:
    names = parameters(layers)
    if isinstance(derivatives, (bool, np.bool_)):
        return names if derivatives else ()
    if isinstance(derivatives, (str, bytes)):
        raise TypeError("derivatives must be a bool or a sequence of parameter names")
    out = tuple(derivatives)
    for name in out:
        if name not in names:
            raise ValueError(f"unknown parameter {name!r}; expected one of {names}")
    return out


def _index(params: tuple[str, ...], param: str, real: str | None = None) -> tuple[int, int]:
    """(layer, part) of ``param``: part 0, 1, 2 for x, Re m, Im m, 3 for the complex d/dm
    (python/scatrs/layered.py:122-136)."""
    match = _NAME.fullmatch(param)
    if match is not None:
        layer, part = int(match[1]), match[2]
        requested = param in params or (
            part == "m" and {f"layer{layer}.m_real", f"layer{layer}.m_imag"} & set(params))
        if requested:
            if part == "m" and real is not None:
                raise ValueError(f"{real} are real: use 'layer{layer}.m_real' or '.m_imag'")
            return layer, ("x", "m_real", "m_imag", "m").index(part)
    raise ValueError(f"no derivative with respect to {param!r}; computed: {params}")


@dataclass(frozen=True, eq=False)
class LayeredEfficiencies(MieEfficiencies):
    """Efficiencies of layered spheres per unit outer cross section; ``d`` takes the layers'
    parameter names."""

    def d(self, param: str) -> MieEfficiencies:
        layer, p = _index(self.params, param, "efficiencies")
        return MieEfficiencies(*self._jacobian[layer, p])


@dataclass(frozen=True, eq=False)
class LayeredCoefficients(MieCoefficients):
    """a_n, b_n of layered spheres; ``d`` takes the layers' parameter names
    ("layer{l}.m": the complex derivative d/dm_l)."""

    def d(self, param: str) -> MieCoefficients:
        layer, p = _index(self.params, param)
        da_dx, da_dm, db_dx, db_dm = self._derivatives[layer]
        if p == 0:
            return MieCoefficients(da_dx, db_dx, self.n_terms)
        if p == 2:
            return MieCoefficients(1j * da_dm, 1j * db_dm, self.n_terms)
        return MieCoefficients(da_dm, db_dm, self.n_terms)


def _device(x, m):
    import torch

    if x.dtype != torch.float64 or m.dtype != torch.complex128:
        raise TypeError("x must be float64 and m complex128")
    if x.device.type != "cuda" or m.device != x.device:
        raise ValueError("x and m must be on the same GPU")
    if x.dim() != 2 or m.shape != x.shape:
        raise ValueError("x and m must have one shape (P, L)")
    if x.shape[1] < 1 or x.shape[1] > MAX_LAYERS:
        raise ValueError(f"layers must be in [1, {MAX_LAYERS}], got {x.shape[1]}")
    return x.contiguous(), m.contiguous()


def efficiencies_device(x, m, *, n_max: int | None = None, derivatives: bool = False,
                        allow_unvalidated: bool = False):
    """Device-resident form: ``x`` float64 (P, L), ``m`` complex128 (P, L) CUDA tensors; returns
    (values (5, P), jacobian (L, 3, 5, P) or None, total terms)."""
    import torch

    x, m = _device(x, m)
    P, L = x.shape
    ctx = _native.context(x.device.index)
    values = torch.empty((5, P), dtype=torch.float64, device=x.device)
    jac = (torch.empty((L, 3, 5, P), dtype=torch.float64, device=x.device)
           if derivatives else None)
    terms = C.c_int64(0)
    ctx.check(ctx.lib.scatrs_layered_efficiencies(
        ctx.handle, C.c_void_p(x.data_ptr()), C.c_void_p(m.data_ptr()), P, L,
        0 if n_max is None else int(n_max), int(bool(allow_unvalidated)),
        C.c_void_p(values.data_ptr()), C.c_void_p(jac.data_ptr()) if jac is not None else None,
        C.byref(terms)))
    return values, jac, terms.value


def coefficients_device(x, m, *, n_max: int | None = None, derivatives: bool = False,
                        allow_unvalidated: bool = False):
    """Device-resident form: returns (n_terms int64 (P,), a, b complex128 (P, N), derivatives
    complex128 (L, 4, P, N) or None)."""
    import torch

    x, m = _device(x, m)
    P, L = x.shape
    ctx = _native.context(x.device.index)
    nm = 0 if n_max is None else int(n_max)
    width, total = C.c_int64(0), C.c_int64(0)
    ctx.check(ctx.lib.scatrs_layered_terms(
        ctx.handle, C.c_void_p(x.data_ptr()), C.c_void_p(m.data_ptr()), P, L, nm,
        int(bool(allow_unvalidated)), int(bool(derivatives)), C.byref(width), C.byref(total)))
    w = max(width.value, 1)
    dev = x.device
    terms = torch.empty(P, dtype=torch.int64, device=dev)
    a = torch.empty((P, w), dtype=torch.complex128, device=dev)
    b = torch.empty((P, w), dtype=torch.complex128, device=dev)
    d = torch.empty((L, 4, P, w), dtype=torch.complex128, device=dev) if derivatives else None
    ctx.check(ctx.lib.scatrs_layered_coefficients(
        ctx.handle, C.c_void_p(x.data_ptr()), C.c_void_p(m.data_ptr()), P, L, nm,
        int(bool(allow_unvalidated)), w, C.c_void_p(terms.data_ptr()), C.c_void_p(a.data_ptr()),
        C.c_void_p(b.data_ptr()), C.c_void_p(d.data_ptr()) if d is not None else None,
        C.byref(total)))
    return terms, a, b, d


def _to_device(xb, mb):
    import torch

    dev = torch.device("cuda", torch.cuda.current_device())
    return torch.from_numpy(xb).to(dev), torch.from_numpy(mb).to(dev)


def efficiencies(x, m, *, n_max: int | None = None, derivatives: bool | Sequence[str] = False,
                 allow_unvalidated: bool = False) -> LayeredEfficiencies:
    """Efficiencies and asymmetry parameter of layered spheres (host arrays in and out)."""
    xb, mb, shape = _inputs(x, m)
    params = _parameters(derivatives, xb.shape[1])
    xt, mt = _to_device(xb, mb)
    values, jac, _ = efficiencies_device(xt, mt, n_max=n_max, derivatives=bool(params),
                                         allow_unvalidated=allow_unvalidated)
    v = values.cpu().numpy()
    j = None if jac is None else jac.cpu().numpy().reshape((*jac.shape[:3], *shape))
    return LayeredEfficiencies(*[v[k].reshape(shape) for k in range(5)], params=params,
                               _jacobian=j)


def coefficients(x, m, *, n_max: int | None = None, derivatives: bool | Sequence[str] = False,
                 allow_unvalidated: bool = False) -> LayeredCoefficients:
    """Mie coefficients a_n and b_n of layered spheres, shape S + (N,) zero-padded to the
    largest number of terms N."""
    xb, mb, shape = _inputs(x, m)
    params = _parameters(derivatives, xb.shape[1])
    xt, mt = _to_device(xb, mb)
    terms, a, b, d = coefficients_device(xt, mt, n_max=n_max, derivatives=bool(params),
                                         allow_unvalidated=allow_unvalidated)
    w = a.shape[1]
    dd = None if d is None else d.cpu().numpy().reshape((*d.shape[:2], *shape, w))
    return LayeredCoefficients(a.cpu().numpy().reshape((*shape, w)),
                               b.cpu().numpy().reshape((*shape, w)),
                               terms.cpu().numpy().reshape(shape), params, dd)
