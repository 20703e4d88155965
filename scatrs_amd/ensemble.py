"""Ensembles of spheres and layered spheres over a size distribution, on the GPU.

Mirrors ``scatrs.Ensemble(particle, size).compute(wavelength, greek=..., normalized=...)``
(python/scatrs/ensemble.py:484-591) and its result layout (SCALARS order,
python/scatrs/ensemble.py:51-63; Greek and phase arrays of ensemble.py:110-140), on the fixed
discrete path of the library: ``n_radius`` Gauss-Legendre nodes in ln r over the
distribution's fixed range and the phase matrix at the ``n_angles`` Gauss-Legendre cosines.
The reference's adaptive Gauss-Kronrod integration (its ``rtol``, ``max_panels``, error
estimates and ``converged`` flags) is not part of this path.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .errors import InvalidInputError
from .particles import LayeredSphere, Sphere
from .size import Gamma, LogNormal, check

__all__ = ["SCALARS", "Ensemble", "EnsembleGreek", "EnsemblePhase", "EnsembleResult",
           "ensemble_fixed", "ensemble_fixed_device", "ensemble_fixed_vjp_device",
           "parameter_names"]

SCALARS = ("c_ext", "c_sca", "c_abs", "c_back", "ssa", "g", "lidar_ratio", "depolarization",
           "effective_radius", "mean_area", "mean_volume")


def ensemble_fixed_device(kind, p0, p1, wavelength, m, fractions=None, *, n_radius: int = 128,
                          n_angles: int = 128, order: int | None = 32, normalized: bool = True,
                          allow_unvalidated: bool = False, want_phase: bool = True):
    """Device-resident batch: every argument a CUDA tensor with one row per ensemble
    (kind int32; p0, p1, wavelength float64; m complex128 (count,) for spheres or (count, L)
    for layered spheres; fractions float64 (count, L-1) for layered spheres). Returns
    (scalars (count, 11), greek (count, 6, order+1) or None, phase (count, 6, n_angles) or
    None, total Mie terms)."""
    import torch

    dev = p0.device
    count = p0.numel()
    layers = 1 if m.dim() == 1 else m.shape[1]
    if layers > 1 and (fractions is None or fractions.shape != (count, layers - 1)):
        raise InvalidInputError("layered spheres need fractions of shape (count, L - 1)")
    args = [kind.to(torch.int32).contiguous(), p0.contiguous(), p1.contiguous(),
            wavelength.contiguous(), m.contiguous()]
    fr = None if layers == 1 else fractions.contiguous()
    scalars = torch.empty((count, 11), dtype=torch.float64, device=dev)
    greek = None if order is None else torch.empty((count, 6, order + 1), dtype=torch.float64,
                                                   device=dev)
    phase = torch.empty((count, 6, n_angles), dtype=torch.float64, device=dev) if want_phase \
        else None
    ctx = _native.context(dev.index)
    terms = C.c_int64(0)
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
    rc = ctx.lib.scatrs_ensemble_fixed(
        ctx.handle, count, p(args[0]), p(args[1]), p(args[2]), p(args[3]), layers, p(fr),
        p(args[4]), int(n_radius), int(n_angles), -1 if order is None else int(order),
        int(bool(normalized)), int(bool(allow_unvalidated)), p(scalars), p(greek), p(phase),
        C.byref(terms))
    ctx.check(rc)
    return scalars, greek, phase, terms.value


def parameter_names(layers: int) -> list[str]:
    """Names of the gradient columns of ensemble_fixed_vjp_device, in the reference's
    parameter vocabulary (scatrs/src/ensemble/mod.rs:182-250): the distribution's two
    parameters, then Re m and Im m per layer (core first), then the radius fractions."""
    names = ["p0", "p1"]
    if layers == 1:
        return names + ["m_real", "m_imag"]
    for l in range(layers):
        names += [f"layer{l}.m_real", f"layer{l}.m_imag"]
    return names + [f"layer{l}.radius_fraction" for l in range(layers - 1)]


def ensemble_fixed_vjp_device(kind, p0, p1, wavelength, m, fractions=None, *, bar_scalars=None,
                              bar_greek=None, bar_phase=None, n_radius: int = 128,
                              n_angles: int = 128, order: int | None = 32,
                              normalized: bool = True, fixed_range: bool = False,
                              allow_unvalidated: bool = False):
    """Reverse mode of ensemble_fixed_device: with the output cotangents bar_scalars
    (count, 11), bar_greek (count, 6, order+1) and bar_phase (count, 6, n_angles) (each may be
    None), the gradient (count, P) with respect to parameter_names(L), and the total Mie
    terms of one pass. fixed_range holds the ln r range at its value (the reference's
    derivative definition); otherwise the radius nodes move with (p0, p1)."""
    import torch

    dev = p0.device
    count = p0.numel()
    layers = 1 if m.dim() == 1 else m.shape[1]
    if layers > 1 and (fractions is None or fractions.shape != (count, layers - 1)):
        raise InvalidInputError("layered spheres need fractions of shape (count, L - 1)")
    if bar_greek is not None and (order is None or bar_greek.shape != (count, 6, order + 1)):
        raise InvalidInputError("bar_greek must have shape (count, 6, order + 1)")
    if bar_scalars is not None and bar_scalars.shape != (count, 11):
        raise InvalidInputError("bar_scalars must have shape (count, 11)")
    if bar_phase is not None and bar_phase.shape != (count, 6, n_angles):
        raise InvalidInputError("bar_phase must have shape (count, 6, n_angles)")
    c = lambda t: None if t is None else t.to(torch.float64).contiguous()  # noqa: E731
    args = [kind.to(torch.int32).contiguous(), p0.contiguous(), p1.contiguous(),
            wavelength.contiguous(), m.contiguous()]
    fr = None if layers == 1 else fractions.contiguous()
    bs, bg, bp = c(bar_scalars), c(bar_greek), c(bar_phase)
    P = 2 + 2 * layers + (layers - 1)
    grad = torch.empty((count, P), dtype=torch.float64, device=dev)
    ctx = _native.context(dev.index)
    terms = C.c_int64(0)
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
    rc = ctx.lib.scatrs_ensemble_fixed_vjp(
        ctx.handle, count, p(args[0]), p(args[1]), p(args[2]), p(args[3]), layers, p(fr),
        p(args[4]), int(n_radius), int(n_angles), -1 if order is None else int(order),
        int(bool(normalized)), int(bool(fixed_range)), int(bool(allow_unvalidated)), p(bs),
        p(bg), p(bp), p(grad), C.byref(terms))
    ctx.check(rc)
    return grad, terms.value


def _autograd_function():
    import torch

    class EnsembleFixedFunction(torch.autograd.Function):
        """(p0, p1, m_real, m_imag, fractions) -> (scalars, greek, phase) on the fixed path;
        backward is the library's reverse-mode kernel (one VJP pass per backward)."""

        @staticmethod
        def forward(ctx, p0, p1, m_real, m_imag, fractions, kind, wavelength, opts):
            ctx.set_materialize_grads(False)  # unused outputs reach backward as None
            m = torch.complex(m_real, m_imag)
            if m.dim() == 2 and m.shape[1] == 1:
                m = m[:, 0]
            sc, gr, ph, _ = ensemble_fixed_device(
                kind, p0, p1, wavelength, m, fractions, n_radius=opts["n_radius"],
                n_angles=opts["n_angles"], order=opts["order"], normalized=opts["normalized"],
                allow_unvalidated=opts["allow_unvalidated"], want_phase=True)
            ctx.save_for_backward(p0, p1, m_real, m_imag,
                                  fractions if fractions is not None else torch.empty(0),
                                  kind, wavelength)
            ctx.opts = opts
            ctx.has_fr = fractions is not None
            return sc, gr, ph

        @staticmethod
        def backward(ctx, g_sc, g_gr, g_ph):
            p0, p1, m_real, m_imag, fr, kind, wl = ctx.saved_tensors
            o = ctx.opts
            m = torch.complex(m_real, m_imag)
            layers = 1 if m.dim() == 1 else m.shape[1]
            if m.dim() == 2 and layers == 1:
                m = m[:, 0]
            grad, _ = ensemble_fixed_vjp_device(
                kind, p0, p1, wl, m, fr if ctx.has_fr else None, bar_scalars=g_sc,
                bar_greek=g_gr, bar_phase=g_ph, n_radius=o["n_radius"], n_angles=o["n_angles"],
                order=o["order"], normalized=o["normalized"], fixed_range=o["fixed_range"],
                allow_unvalidated=o["allow_unvalidated"])
            gm_re = grad[:, 2:2 + 2 * layers:2].reshape(m_real.shape)
            gm_im = grad[:, 3:3 + 2 * layers:2].reshape(m_imag.shape)
            gfr = grad[:, 2 + 2 * layers:] if ctx.has_fr else None
            return grad[:, 0], grad[:, 1], gm_re, gm_im, gfr, None, None, None

    return EnsembleFixedFunction


_FUNCTION = None


def ensemble_fixed(kind, p0, p1, wavelength, m_real, m_imag, fractions=None, *,
                   n_radius: int = 128, n_angles: int = 128, order: int = 32,
                   normalized: bool = True, fixed_range: bool = False,
                   allow_unvalidated: bool = False):
    """Differentiable batch of ensembles on the fixed path (torch.autograd): p0, p1,
    wavelength (count,); m_real, m_imag (count,) for spheres or (count, L) for layered spheres
    (core first); fractions (count, L-1). Returns (scalars (count, 11), greek
    (count, 6, order+1), phase (count, 6, n_angles)); gradients flow to p0, p1, m_real,
    m_imag and fractions through the library's reverse-mode kernel."""
    global _FUNCTION
    if _FUNCTION is None:
        _FUNCTION = _autograd_function()
    opts = dict(n_radius=int(n_radius), n_angles=int(n_angles), order=int(order),
                normalized=bool(normalized), fixed_range=bool(fixed_range),
                allow_unvalidated=bool(allow_unvalidated))
    return _FUNCTION.apply(p0, p1, m_real, m_imag, fractions, kind, wavelength, opts)


@dataclass(frozen=True, eq=False)
class EnsembleGreek:
    """Greek coefficients, arrays of shape grid + wavelength.shape + (L + 1,)."""

    alpha1: np.ndarray
    alpha2: np.ndarray
    alpha3: np.ndarray
    alpha4: np.ndarray
    beta1: np.ndarray
    beta2: np.ndarray


@dataclass(frozen=True, eq=False)
class EnsemblePhase:
    """The phase matrix at the cosines ``mu`` (the n_angles Gauss-Legendre nodes), arrays of
    shape grid + wavelength.shape + (n_angles,); f22 = f11 and f44 = f33 for spheres."""

    mu: np.ndarray
    f11: np.ndarray
    f12: np.ndarray
    f22: np.ndarray
    f33: np.ndarray
    f34: np.ndarray
    f44: np.ndarray


@dataclass(frozen=True, eq=False)
class EnsembleResult:
    """Outputs of an ensemble, each of shape grid + wavelength.shape (SCALARS order)."""

    c_ext: np.ndarray
    c_sca: np.ndarray
    c_abs: np.ndarray
    c_back: np.ndarray
    ssa: np.ndarray
    g: np.ndarray
    lidar_ratio: np.ndarray
    depolarization: np.ndarray
    effective_radius: np.ndarray
    mean_area: np.ndarray
    mean_volume: np.ndarray
    greek: EnsembleGreek | None = None
    phase: EnsemblePhase | None = None
    wavelength: np.ndarray = field(default=None)
    terms: int = 0


def _frozen(a: np.ndarray) -> np.ndarray:
    a.flags.writeable = False
    return a


@dataclass(frozen=True)
class Ensemble:
    """Particles of one kind over a size distribution (python/scatrs/ensemble.py:484-529).

    ``size`` may have array-valued parameters: a grid of distributions, computed at once
    (the results have the grid's axes first)."""

    particle: Sphere | LayeredSphere
    size: LogNormal | Gamma

    def __post_init__(self) -> None:
        if not isinstance(self.particle, (Sphere, LayeredSphere)):
            raise TypeError("particle must be a Sphere or a LayeredSphere, not "
                            f"{type(self.particle).__name__}")
        check(self.size)

    def compute(self, wavelength, *, greek: int | None = None, normalized: bool = True,
                n_radius: int = 128, n_angles: int = 128, allow_unvalidated: bool = False,
                phase: bool = True) -> EnsembleResult:
        """Averages over the ensemble at each wavelength on the fixed discrete path."""
        import torch

        wl = np.asarray(wavelength, np.float64)
        p0, p1 = self.size.params()
        grid = np.broadcast(p0, p1).shape
        p0g, p1g = (np.broadcast_to(v, grid).reshape(-1) for v in (p0, p1))
        nd, nw = p0g.size, wl.size
        m = self.particle.indices(wl).reshape(self.particle.layers, -1)  # (L, nw)
        L = self.particle.layers
        dev = torch.device("cuda", torch.cuda.current_device())
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        kind = np.full(nd * nw, self.size.kind, np.int32)
        mm = np.tile(m.T, (nd, 1))  # (nd * nw, L)
        fr = None
        if L > 1:
            fr = t(np.tile(self.particle.fractions(), (nd * nw, 1)))
        sc, gr, ph, terms = ensemble_fixed_device(
            t(kind), t(np.repeat(p0g, nw)), t(np.repeat(p1g, nw)), t(np.tile(wl.reshape(-1), nd)),
            t(mm[:, 0] if L == 1 else mm), fr, n_radius=n_radius, n_angles=n_angles,
            order=greek, normalized=normalized, allow_unvalidated=allow_unvalidated,
            want_phase=phase)
        shape = grid + wl.shape
        s = sc.cpu().numpy()
        fields = [_frozen(s[:, k].reshape(shape)) for k in range(11)]
        g = None
        if gr is not None:
            a = gr.cpu().numpy()
            g = EnsembleGreek(*(_frozen(a[:, k].reshape(shape + (a.shape[-1],)))
                                for k in range(6)))
        p = None
        if ph is not None:
            from .quadrature import gauss_legendre

            a = ph.cpu().numpy()
            mu, _ = gauss_legendre(n_angles)
            p = EnsemblePhase(_frozen(mu), *(_frozen(a[:, k].reshape(shape + (n_angles,)))
                                             for k in range(6)))
        return EnsembleResult(*fields, greek=g, phase=p, wavelength=_frozen(wl.copy()),
                              terms=terms)
