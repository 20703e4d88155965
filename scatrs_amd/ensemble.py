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
import math
from dataclasses import dataclass, field

import numpy as np

from . import _native
from .errors import InvalidInputError
from .particles import LayeredSphere, Sphere
from .size import SEGMENT, Distribution, MultiModal, Tabulated, check

__all__ = ["SCALARS", "Ensemble", "EnsembleGraph", "EnsembleGreek", "EnsemblePhase", "EnsembleResult",
           "Mixture", "ensemble_fixed", "ensemble_fixed_device", "ensemble_fixed_vjp_device",
           "ensemble_nodes", "ensemble_nodes_device", "ensemble_nodes_tangent_device",
           "ensemble_nodes_vjp_device",
           "assemble", "mix_parts", "mix_units", "number_averages", "parameter_names"]

SCALARS = ("c_ext", "c_sca", "c_abs", "c_back", "ssa", "g", "lidar_ratio", "depolarization",
           "effective_radius", "mean_area", "mean_volume")


def ensemble_fixed_device(kind, p0, p1, wavelength, m, fractions=None, *, p2=None,
                          n_radius: int = 128, n_angles: int = 128, order: int | None = 32,
                          normalized: bool = True, allow_unvalidated: bool = False,
                          want_phase: bool = True, mu=None):
    """Device-resident batch: every argument a CUDA tensor with one row per ensemble
    (kind int32; p0, p1, wavelength float64; p2 float64 or None, the third parameter of a
    modified gamma or power law; m complex128 (count,) for spheres or (count, L) for layered
    spheres; fractions float64 (count, L-1) for layered spheres). Returns (scalars (count, 11),
    greek (count, 6, order+1) or None, phase (count, 6, n_angles) or None, total Mie
    terms); with ``mu`` (host cosines in [-1, 1]) the phase matrix is (count, 6, len(mu)) at
    those cosines (the reference's mu argument)."""
    import torch

    dev = p0.device
    count = p0.numel()
    layers = 1 if m.dim() == 1 else m.shape[1]
    if layers > 1 and (fractions is None or fractions.shape != (count, layers - 1)):
        raise InvalidInputError("layered spheres need fractions of shape (count, L - 1)")
    args = [kind.to(torch.int32).contiguous(), p0.contiguous(), p1.contiguous(),
            wavelength.contiguous(), m.contiguous()]
    fr = None if layers == 1 else fractions.contiguous()
    q2 = None if p2 is None else p2.to(torch.float64).contiguous()
    scalars = torch.empty((count, 11), dtype=torch.float64, device=dev)
    greek = None if order is None else torch.empty((count, 6, order + 1), dtype=torch.float64,
                                                   device=dev)
    n_phase = n_angles if mu is None else int(np.size(mu))
    phase = torch.empty((count, 6, n_phase), dtype=torch.float64, device=dev) if want_phase \
        else None
    ctx = _native.context(dev.index)
    terms = C.c_int64(0)
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
    if mu is not None:
        ctx.set_phase_cosines(mu)
    try:
        rc = ctx.lib.scatrs_ensemble_fixed(
            ctx.handle, count, p(args[0]), p(args[1]), p(args[2]), p(q2), p(args[3]), layers,
            p(fr), p(args[4]), int(n_radius), int(n_angles), -1 if order is None else int(order),
            int(bool(normalized)), int(bool(allow_unvalidated)), p(scalars), p(greek), p(phase),
            C.byref(terms))
    finally:
        if mu is not None:
            ctx.set_phase_cosines(None)
    ctx.check(rc)
    return scalars, greek, phase, terms.value


class EnsembleGraph:
    """A forward ensemble batch captured as a HIP graph (scatrs_ensemble_fixed_graph): made
    from the arguments of ensemble_fixed_device, whose outputs it computes once; ``launch()``
    recomputes them into the same output tensors by replaying the captured kernels (no host
    synchronization, no per-kernel launch cost), reading the same input tensors, which may be
    changed in place as long as no radius node changes its number of terms (``status()``
    reports it). The graph keeps its inputs and outputs alive."""

    def __init__(self, kind, p0, p1, wavelength, m, fractions=None, *, p2=None,
                 n_radius: int = 128, n_angles: int = 128, order: int | None = 32,
                 normalized: bool = True, allow_unvalidated: bool = False,
                 want_phase: bool = True):
        import torch

        dev = p0.device
        count = p0.numel()
        layers = 1 if m.dim() == 1 else m.shape[1]
        if layers > 1 and (fractions is None or fractions.shape != (count, layers - 1)):
            raise InvalidInputError("layered spheres need fractions of shape (count, L - 1)")
        self.inputs = [kind.to(torch.int32).contiguous(), p0.contiguous(), p1.contiguous(),
                       wavelength.contiguous(), m.contiguous(),
                       None if layers == 1 else fractions.contiguous(),
                       None if p2 is None else p2.to(torch.float64).contiguous()]
        self.scalars = torch.empty((count, 11), dtype=torch.float64, device=dev)
        self.greek = None if order is None else torch.empty((count, 6, order + 1),
                                                            dtype=torch.float64, device=dev)
        self.phase = (torch.empty((count, 6, n_angles), dtype=torch.float64, device=dev)
                      if want_phase else None)
        ptr = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
        self.ctx = _native.context(dev.index)
        k, q0, q1, wl, mm, fr, q2 = self.inputs
        terms = C.c_int64(0)
        handle = C.c_void_p()
        self.ctx.check(self.ctx.lib.scatrs_ensemble_fixed_graph(
            self.ctx.handle, count, ptr(k), ptr(q0), ptr(q1), ptr(q2), ptr(wl), layers, ptr(fr),
            ptr(mm), n_radius, n_angles, -1 if order is None else order, int(bool(normalized)),
            int(bool(allow_unvalidated)), ptr(self.scalars), ptr(self.greek), ptr(self.phase),
            C.byref(terms), C.byref(handle)))
        self.handle = handle
        self.terms = terms.value
        self.lib = self.ctx.lib

    def launch(self, stream=None) -> None:
        """Replays the call on ``stream`` (default: torch's current stream)."""
        import torch

        s = torch.cuda.current_stream(self.scalars.device) if stream is None else stream
        rc = self.lib.scatrs_graph_launch(self.handle, C.c_void_p(s.cuda_stream))
        if rc != _native.OK:
            raise _native.NativeLibraryError(self.lib.scatrs_graph_last_error(self.handle).decode())

    def status(self, stream=None) -> None:
        """Waits for the replays on ``stream`` and raises InvalidInputError if an input changed a
        node's number of terms since the capture."""
        import torch

        s = torch.cuda.current_stream(self.scalars.device) if stream is None else stream
        rc = self.lib.scatrs_graph_status(self.handle, C.c_void_p(s.cuda_stream))
        if rc == _native.INVALID_INPUT:
            raise InvalidInputError(self.lib.scatrs_graph_last_error(self.handle).decode())
        if rc != _native.OK:
            raise _native.NativeLibraryError(self.lib.scatrs_graph_last_error(self.handle).decode())

    def close(self) -> None:
        if getattr(self, "handle", None):
            self.lib.scatrs_graph_destroy(self.handle)
            self.handle = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:
            pass


def ensemble_nodes_device(nodes, wavelength, m, fractions=None, *, n_angles: int = 128,
                          order: int | None = 32, normalized: bool = True,
                          allow_unvalidated: bool = False, want_phase: bool = True, mu=None):
    """Ensembles on the caller's radius nodes (scatrs_ensemble_nodes): nodes float64 (count,
    n_nodes, 2) of (r, W) pairs, W the quadrature weight times the number density per ln r;
    wavelength, m, fractions and the outputs as ensemble_fixed_device's."""
    import torch

    dev = nodes.device
    count, n_nodes = nodes.shape[0], nodes.shape[1]
    layers = 1 if m.dim() == 1 else m.shape[1]
    if nodes.shape != (count, n_nodes, 2):
        raise InvalidInputError("nodes must have shape (count, n_nodes, 2)")
    if layers > 1 and (fractions is None or fractions.shape != (count, layers - 1)):
        raise InvalidInputError("layered spheres need fractions of shape (count, L - 1)")
    nd = nodes.to(torch.float64).contiguous()
    wl, mm = wavelength.contiguous(), m.contiguous()
    fr = None if layers == 1 else fractions.contiguous()
    scalars = torch.empty((count, 11), dtype=torch.float64, device=dev)
    greek = None if order is None else torch.empty((count, 6, order + 1), dtype=torch.float64,
                                                   device=dev)
    n_phase = n_angles if mu is None else int(np.size(mu))
    phase = torch.empty((count, 6, n_phase), dtype=torch.float64, device=dev) if want_phase \
        else None
    ctx = _native.context(dev.index)
    terms = C.c_int64(0)
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
    if mu is not None:
        ctx.set_phase_cosines(mu)
    try:
        rc = ctx.lib.scatrs_ensemble_nodes(
            ctx.handle, count, n_nodes, p(nd), p(wl), layers, p(fr), p(mm), int(n_angles),
            -1 if order is None else int(order), int(bool(normalized)),
            int(bool(allow_unvalidated)), p(scalars), p(greek), p(phase), C.byref(terms))
    finally:
        if mu is not None:
            ctx.set_phase_cosines(None)
    ctx.check(rc)
    return scalars, greek, phase, terms.value


def ensemble_nodes_tangent_device(nodes, wavelength, m, fractions=None, *, direction: int,
                                  sign: int, n_angles: int = 128, order: int | None = 32,
                                  allow_unvalidated: bool = False, want_phase: bool = True,
                                  mu=None):
    """scatrs_ensemble_nodes_tangent: ensemble_nodes_device (normalized=False) with every node's
    coefficients moved by sign eps dc along the particle's ``direction`` (2 l, 2 l + 1: Re m,
    Im m of layer l; 2 L + j: radius fraction j), eps = 1/(1 + x), and its weight scaled by
    1/(2 eps). Returns (scalars, greek, phase, terms): averages over the scaled weights."""
    import torch

    dev = nodes.device
    count, n_nodes = nodes.shape[0], nodes.shape[1]
    layers = 1 if m.dim() == 1 else m.shape[1]
    nd = nodes.to(torch.float64).contiguous()
    wl, mm = wavelength.contiguous(), m.contiguous()
    fr = None if layers == 1 else fractions.contiguous()
    scalars = torch.empty((count, 11), dtype=torch.float64, device=dev)
    greek = None if order is None else torch.empty((count, 6, order + 1), dtype=torch.float64,
                                                   device=dev)
    n_phase = n_angles if mu is None else int(np.size(mu))
    phase = torch.empty((count, 6, n_phase), dtype=torch.float64, device=dev) if want_phase \
        else None
    ctx = _native.context(dev.index)
    terms = C.c_int64(0)
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
    if mu is not None:
        ctx.set_phase_cosines(mu)
    try:
        rc = ctx.lib.scatrs_ensemble_nodes_tangent(
            ctx.handle, count, n_nodes, p(nd), p(wl), layers, p(fr), p(mm), int(n_angles),
            -1 if order is None else int(order), int(direction), int(sign),
            int(bool(allow_unvalidated)), p(scalars), p(greek), p(phase), C.byref(terms))
    finally:
        if mu is not None:
            ctx.set_phase_cosines(None)
    ctx.check(rc)
    return scalars, greek, phase, terms.value


def _scalars_tangent(avg, davg):
    """The tangent of assemble's scalars (normalized=False) at the averages avg (..., 7) along
    davg: d(a / b) = (da - (a / b) db) / b for the ratios, depolarization 0."""
    import torch

    ext, sca, absb, back, scag, r2, r3 = avg.unbind(-1)
    de, ds, da, db, dg, d2, d3 = davg.unbind(-1)
    q = lambda a, b, dta, dtb: (dta - (a / b) * dtb) / b  # noqa: E731
    return torch.stack([de, ds, da, db, q(sca, ext, ds, de), q(scag, sca, dg, ds),
                        4.0 * math.pi * q(ext, back, de, db), torch.zeros_like(de),
                        q(r3, r2, d3, d2), math.pi * d2, 4.0 / 3.0 * math.pi * d3], dim=-1)


def _nodes_jvp(o, sc, gr, ph, m_real, m_imag, fractions, nodes, dlnw, wavelength, dp, dm_re,
               dm_im, dfr):
    """Tangents of (scalars, greek, phase) of ensemble_nodes (normalized=False; rows in runs of
    o["replicas"] = k equal replicas, each with its own tangent) along the input tangents.

    The weights' part: one more call on the rows whose d ln W . dp is not zero, with the weights
    W (s + c) > 0, s = d ln W . dp, c = 1 + max |s| per row; its sums minus c times the primal's
    are sum W s F. Each particle direction's part (2 l, 2 l + 1: Re m, Im m of layer l; 2 L + j:
    radius fraction j): two tangent calls (scatrs_ensemble_nodes_tangent, sign +1 and -1) on one
    row of every run with a nonzero tangent, whose sums' difference is sum W dF per unit step,
    times each row's tangent. Averages move by (dS - A dZ) / Z."""
    import torch

    k = o["replicas"]
    W = nodes[..., 1]
    z = W.sum(1)
    m = torch.complex(m_real, m_imag)
    layers = 1 if m.dim() == 1 else m.shape[1]
    fr = fractions if layers > 1 else None
    W6 = 6 * (gr.shape[-1] if gr is not None else 0)
    nph = ph.shape[-1] if ph is not None else 0
    sub = lambda a, rows: None if a is None else a[rows]  # noqa: E731

    def flat(s, g, p):
        cols = [number_averages(s)]
        if g is not None:
            cols.append(g.reshape(s.shape[0], -1))
        if p is not None:
            cols.append(p.reshape(s.shape[0], -1))
        return torch.cat(cols, 1)

    A = flat(sc, gr, ph)  # averages of the linear quantities over W
    dS = torch.zeros_like(A)  # tangents of their weighted sums
    dZ = torch.zeros_like(z)
    kw = dict(n_angles=o["n_angles"], order=o["order"], allow_unvalidated=o["allow_unvalidated"],
              want_phase=o["phase"], mu=o["mu"])
    if dp is not None and dlnw is not None:
        sw = (dlnw * dp[:, None, :]).sum(-1)
        rows = torch.nonzero((sw != 0).any(1)).flatten()
        if rows.numel():
            s_r, w_r = sw[rows], W[rows]
            c = s_r.abs().amax(1) + 1.0
            ww = w_r * (s_r + c[:, None])
            nw = torch.stack([nodes[rows, :, 0], ww], -1)
            s2, g2, p2, _ = ensemble_nodes_device(nw, wavelength[rows], m[rows], sub(fr, rows),
                                                  normalized=False, **kw)
            dS[rows] += flat(s2, g2, p2) * ww.sum(1)[:, None] - c[:, None] * A[rows] * \
                z[rows, None]
            dZ[rows] += (w_r * s_r).sum(1)
    dirs = []
    for l in range(layers):
        for part, dm in ((0, dm_re), (1, dm_im)):
            if dm is not None:
                dirs.append((2 * l + part, dm if dm.dim() == 1 else dm[:, l]))
    if dfr is not None and layers > 1:
        for j in range(layers - 1):
            dirs.append((2 * layers + j, dfr[:, j]))
    x = 2.0 * math.pi * nodes[..., 0] / wavelength[:, None]
    zs = (W * (1.0 + x) * 0.5).sum(1)  # the tangent calls' sums of weights
    span = torch.arange(k, device=nodes.device)
    for d, t in dirs:
        runs = torch.nonzero((t != 0).reshape(-1, k).any(1)).flatten()
        if not runs.numel():
            continue
        u = runs * k  # a run's rows share their nodes and particle
        one = lambda sign: flat(*ensemble_nodes_tangent_device(  # noqa: E731
            nodes[u], wavelength[u], m[u], sub(fr, u), direction=d, sign=sign, **kw)[:3])
        dd = (one(1) - one(-1)) * zs[u, None]
        rr = (u[:, None] + span[None, :]).reshape(-1)
        dS[rr] += t[rr, None] * dd.repeat_interleave(k, 0)
    dA = (dS - A * dZ[:, None]) / z[:, None]
    dsc = _scalars_tangent(A[:, :7], dA[:, :7])
    dgr = dA[:, 7:7 + W6].reshape(gr.shape).contiguous() if gr is not None else None
    dph = dA[:, 7 + W6:7 + W6 + 6 * nph].reshape(ph.shape).contiguous() if ph is not None \
        else None
    return dsc, dgr, dph


def ensemble_nodes_vjp_device(nodes, dlnw, wavelength, m, fractions=None, *, bar_scalars=None,
                              bar_greek=None, bar_phase=None, n_angles: int = 128,
                              order: int | None = 32, normalized: bool = True,
                              allow_unvalidated: bool = False, mu=None):
    """Reverse mode of ensemble_nodes_device (scatrs_ensemble_nodes_vjp): dlnw (count, n_nodes,
    3) the nodes' d ln W / d p, or None; the cotangents as ensemble_fixed_vjp_device's. Returns
    the gradient (count, P) in parameter_names(L) order and the terms of one pass."""
    import torch

    dev = nodes.device
    count, n_nodes = nodes.shape[0], nodes.shape[1]
    layers = 1 if m.dim() == 1 else m.shape[1]
    if layers > 1 and (fractions is None or fractions.shape != (count, layers - 1)):
        raise InvalidInputError("layered spheres need fractions of shape (count, L - 1)")
    if dlnw is not None and dlnw.shape != (count, n_nodes, 3):
        raise InvalidInputError("dlnw must have shape (count, n_nodes, 3)")
    c = lambda t: None if t is None else t.to(torch.float64).contiguous()  # noqa: E731
    nd, dl = c(nodes), c(dlnw)
    bs, bg, bp = c(bar_scalars), c(bar_greek), c(bar_phase)
    fr = None if layers == 1 else fractions.contiguous()
    P = 3 + 2 * layers + (layers - 1)
    grad = torch.empty((count, P), dtype=torch.float64, device=dev)
    ctx = _native.context(dev.index)
    terms = C.c_int64(0)
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
    if mu is not None:
        ctx.set_phase_cosines(mu)
    try:
        rc = ctx.lib.scatrs_ensemble_nodes_vjp(
            ctx.handle, count, n_nodes, p(nd), p(dl), p(wavelength.contiguous()), layers, p(fr),
            p(m.contiguous()), int(n_angles), -1 if order is None else int(order),
            int(bool(normalized)), int(bool(allow_unvalidated)), p(bs), p(bg), p(bp), p(grad),
            C.byref(terms))
    finally:
        if mu is not None:
            ctx.set_phase_cosines(None)
    ctx.check(rc)
    return grad, terms.value


def _nodes_function():
    import torch

    class EnsembleNodesFunction(torch.autograd.Function):
        """(p (count, 3), m_real, m_imag, fractions) -> (scalars, greek, phase) on the caller's
        nodes; the parameters act through the nodes' ln W (derivatives dlnw), the nodes stay."""

        @staticmethod
        def forward(ctx, p, m_real, m_imag, fractions, nodes, dlnw, wavelength, opts):
            ctx.set_materialize_grads(False)
            m = torch.complex(m_real, m_imag)
            k = opts["replicas"]  # rows come in runs of k equal replicas: solve each run once
            sc, gr, ph, _ = ensemble_nodes_device(
                nodes[::k], wavelength[::k], m[::k], None if fractions is None else
                fractions[::k], n_angles=opts["n_angles"], order=opts["order"],
                normalized=opts["normalized"], allow_unvalidated=opts["allow_unvalidated"],
                want_phase=opts["phase"], mu=opts["mu"])
            if k > 1:
                rep = lambda a: None if a is None else a.repeat_interleave(k, 0)  # noqa: E731
                sc, gr, ph = rep(sc), rep(gr), rep(ph)
            empty = torch.empty(0)
            ctx.save_for_backward(m_real, m_imag, fractions if fractions is not None else empty,
                                  nodes, dlnw, wavelength)
            ctx.save_for_forward(m_real, m_imag, fractions if fractions is not None else empty,
                                 nodes, dlnw if dlnw is not None else empty, wavelength)
            ctx.fwd_out = (sc, gr, ph)
            ctx.opts = opts
            ctx.has_fr = fractions is not None
            ctx.has_dlnw = dlnw is not None
            return sc, gr, ph

        @staticmethod
        def jvp(ctx, dp, dm_re, dm_im, dfr, dnodes, ddlnw, dwl, dopts):
            m_real, m_imag, fr, nodes, dlnw, wl = ctx.saved_tensors
            o = ctx.opts
            if o["normalized"]:
                raise NotImplementedError("forward mode needs normalized=False")
            sc, gr, ph = ctx.fwd_out
            return _nodes_jvp(o, sc, gr, ph, m_real, m_imag, fr if ctx.has_fr else None, nodes,
                              dlnw if ctx.has_dlnw else None, wl, dp, dm_re, dm_im,
                              dfr if ctx.has_fr else None)

        @staticmethod
        def backward(ctx, g_sc, g_gr, g_ph):
            m_real, m_imag, fr, nodes, dlnw, wl = ctx.saved_tensors
            o = ctx.opts
            m = torch.complex(m_real, m_imag)
            layers = 1 if m.dim() == 1 else m.shape[1]
            grad, _ = ensemble_nodes_vjp_device(
                nodes, dlnw, wl, m, fr if ctx.has_fr else None, bar_scalars=g_sc, bar_greek=g_gr,
                bar_phase=g_ph, n_angles=o["n_angles"], order=o["order"],
                normalized=o["normalized"], allow_unvalidated=o["allow_unvalidated"],
                mu=o["mu"])
            gm_re = grad[:, 3:3 + 2 * layers:2].reshape(m_real.shape)
            gm_im = grad[:, 4:4 + 2 * layers:2].reshape(m_imag.shape)
            gfr = grad[:, 3 + 2 * layers:] if ctx.has_fr else None
            return grad[:, :3], gm_re, gm_im, gfr, None, None, None, None

    return EnsembleNodesFunction


_NODES_FUNCTION = None


def ensemble_nodes(p, m_real, m_imag, fractions, nodes, dlnw, wavelength, *, n_angles: int = 128,
                   order: int | None = 32, normalized: bool = True,
                   allow_unvalidated: bool = False, mu=None, phase: bool = True,
                   replicas: int = 1):
    """Differentiable ensembles on the caller's nodes (torch.autograd): p (count, 3) the
    parameters the nodes' weights depend on, through dlnw (count, n_nodes, 3) = d ln W / d p
    (the nodes do not move); m_real, m_imag (count,) or (count, L); fractions (count, L-1) or
    None; nodes (count, n_nodes, 2). Returns (scalars, greek, phase) as ensemble_nodes_device.
    With ``replicas`` = k, the rows come in runs of k equal copies (the same nodes and
    parameter values, each its own leaf for its own cotangent): the forward pass solves each
    run once."""
    global _NODES_FUNCTION
    if _NODES_FUNCTION is None:
        _NODES_FUNCTION = _nodes_function()
    opts = dict(n_angles=int(n_angles), order=None if order is None else int(order),
                normalized=bool(normalized), allow_unvalidated=bool(allow_unvalidated),
                phase=bool(phase), mu=None if mu is None else np.array(mu, np.float64).reshape(-1),
                replicas=int(replicas))
    return _NODES_FUNCTION.apply(p, m_real, m_imag, fractions, nodes, dlnw, wavelength, opts)


def parameter_names(layers: int) -> list[str]:
    """Names of the gradient columns of ensemble_fixed_vjp_device, in the reference's
    parameter vocabulary (scatrs/src/ensemble/mod.rs:182-250): the distribution's three
    parameter slots (0 where a kind has no such parameter), then Re m and Im m per layer (core
    first), then the radius fractions."""
    names = ["p0", "p1", "p2"]
    if layers == 1:
        return names + ["m_real", "m_imag"]
    for l in range(layers):
        names += [f"layer{l}.m_real", f"layer{l}.m_imag"]
    return names + [f"layer{l}.radius_fraction" for l in range(layers - 1)]


def ensemble_fixed_vjp_device(kind, p0, p1, wavelength, m, fractions=None, *, p2=None,
                              bar_scalars=None, bar_greek=None, bar_phase=None,
                              n_radius: int = 128,
                              n_angles: int = 128, order: int | None = 32,
                              normalized: bool = True, fixed_range: bool = False,
                              allow_unvalidated: bool = False, mu=None):
    """Reverse mode of ensemble_fixed_device: with the output cotangents bar_scalars
    (count, 11), bar_greek (count, 6, order+1) and bar_phase (count, 6, n_angles) (each may be
    None), the gradient (count, P) with respect to parameter_names(L), and the total Mie
    terms of one pass. fixed_range holds the ln r range at its value (the reference's
    derivative definition); otherwise the radius nodes move with the parameters. With ``mu``
    (host cosines) bar_phase is (count, 6, len(mu)) at those cosines."""
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
    n_phase = n_angles if mu is None else int(np.size(mu))
    if bar_phase is not None and bar_phase.shape != (count, 6, n_phase):
        raise InvalidInputError("bar_phase must have shape (count, 6, n_angles), or "
                                "(count, 6, len(mu)) with mu")
    c = lambda t: None if t is None else t.to(torch.float64).contiguous()  # noqa: E731
    args = [kind.to(torch.int32).contiguous(), p0.contiguous(), p1.contiguous(),
            wavelength.contiguous(), m.contiguous()]
    fr = None if layers == 1 else fractions.contiguous()
    bs, bg, bp = c(bar_scalars), c(bar_greek), c(bar_phase)
    q2 = c(p2)
    P = 3 + 2 * layers + (layers - 1)
    grad = torch.empty((count, P), dtype=torch.float64, device=dev)
    ctx = _native.context(dev.index)
    terms = C.c_int64(0)
    p = lambda t: None if t is None else C.c_void_p(t.data_ptr())  # noqa: E731
    if mu is not None:
        ctx.set_phase_cosines(mu)
    try:
        rc = ctx.lib.scatrs_ensemble_fixed_vjp(
            ctx.handle, count, p(args[0]), p(args[1]), p(args[2]), p(q2), p(args[3]), layers,
            p(fr), p(args[4]), int(n_radius), int(n_angles), -1 if order is None else int(order),
            int(bool(normalized)), int(bool(fixed_range)), int(bool(allow_unvalidated)), p(bs),
            p(bg), p(bp), p(grad), C.byref(terms))
    finally:
        if mu is not None:
            ctx.set_phase_cosines(None)
    ctx.check(rc)
    return grad, terms.value


def _autograd_function():
    import torch

    class EnsembleFixedFunction(torch.autograd.Function):
        """(p0, p1, p2, m_real, m_imag, fractions) -> (scalars, greek, phase) on the fixed
        path; backward is the library's reverse-mode kernel (one VJP pass per backward)."""

        @staticmethod
        def forward(ctx, p0, p1, p2, m_real, m_imag, fractions, kind, wavelength, opts):
            ctx.set_materialize_grads(False)  # unused outputs reach backward as None
            m = torch.complex(m_real, m_imag)
            if m.dim() == 2 and m.shape[1] == 1:
                m = m[:, 0]
            sc, gr, ph, _ = ensemble_fixed_device(
                kind, p0, p1, wavelength, m, fractions, p2=p2, n_radius=opts["n_radius"],
                n_angles=opts["n_angles"], order=opts["order"], normalized=opts["normalized"],
                allow_unvalidated=opts["allow_unvalidated"], want_phase=opts["phase"],
                mu=opts["mu"])
            empty = torch.empty(0)
            ctx.save_for_backward(p0, p1, p2 if p2 is not None else empty, m_real, m_imag,
                                  fractions if fractions is not None else empty, kind, wavelength)
            ctx.opts = opts
            ctx.has_fr = fractions is not None
            ctx.has_p2 = p2 is not None
            return sc, gr, ph

        @staticmethod
        def backward(ctx, g_sc, g_gr, g_ph):
            p0, p1, p2, m_real, m_imag, fr, kind, wl = ctx.saved_tensors
            o = ctx.opts
            m = torch.complex(m_real, m_imag)
            layers = 1 if m.dim() == 1 else m.shape[1]
            if m.dim() == 2 and layers == 1:
                m = m[:, 0]
            grad, _ = ensemble_fixed_vjp_device(
                kind, p0, p1, wl, m, fr if ctx.has_fr else None, p2=p2 if ctx.has_p2 else None,
                bar_scalars=g_sc, bar_greek=g_gr, bar_phase=g_ph, n_radius=o["n_radius"],
                n_angles=o["n_angles"], order=o["order"], normalized=o["normalized"],
                fixed_range=o["fixed_range"], allow_unvalidated=o["allow_unvalidated"],
                mu=o["mu"])
            gm_re = grad[:, 3:3 + 2 * layers:2].reshape(m_real.shape)
            gm_im = grad[:, 4:4 + 2 * layers:2].reshape(m_imag.shape)
            gfr = grad[:, 3 + 2 * layers:] if ctx.has_fr else None
            gp2 = grad[:, 2] if ctx.has_p2 else None
            return grad[:, 0], grad[:, 1], gp2, gm_re, gm_im, gfr, None, None, None

    return EnsembleFixedFunction


_FUNCTION = None


def ensemble_fixed(kind, p0, p1, wavelength, m_real, m_imag, fractions=None, *, p2=None,
                   n_radius: int = 128, n_angles: int = 128, order: int | None = 32,
                   normalized: bool = True, fixed_range: bool = False,
                   allow_unvalidated: bool = False, mu=None, phase: bool = True):
    """Differentiable batch of ensembles on the fixed path (torch.autograd): p0, p1,
    wavelength (count,) and p2 (count,) or None; m_real, m_imag (count,) for spheres or
    (count, L) for layered spheres (core first); fractions (count, L-1). Returns (scalars
    (count, 11), greek (count, 6, order+1) or None without an order, phase (count, 6,
    n_angles), or at the host cosines ``mu`` (count, 6, len(mu)), or None with phase=False);
    gradients flow to p0, p1, p2, m_real, m_imag and fractions through the library's
    reverse-mode kernel."""
    global _FUNCTION
    if _FUNCTION is None:
        _FUNCTION = _autograd_function()
    opts = dict(n_radius=int(n_radius), n_angles=int(n_angles),
                order=None if order is None else int(order),
                normalized=bool(normalized), fixed_range=bool(fixed_range),
                allow_unvalidated=bool(allow_unvalidated), phase=bool(phase),
                mu=None if mu is None else np.array(mu, np.float64).reshape(-1))
    return _FUNCTION.apply(p0, p1, p2, m_real, m_imag, fractions, kind, wavelength, opts)


class SegmentPlan:
    """The layout segment_sum sums by, for one index vector (``index``, on the device): the
    rows' stable order by index, the sorted indices, the distinct indices, the longest segment,
    and each sorted row's segment and place in it."""

    __slots__ = ("index", "order", "ks", "uk", "width", "seg", "pos")

    def __init__(self, index, order, ks, uk, width, seg, pos):
        self.index, self.order, self.ks, self.uk = index, order, ks, uk
        self.width, self.seg, self.pos = width, seg, pos


def segment_plan(index, device):
    """The SegmentPlan of ``index``: from a host array (numpy) without any device
    synchronization (its five arrays go to ``device`` in one transfer), or from a tensor."""
    import torch

    if isinstance(index, np.ndarray):
        idx = np.ascontiguousarray(index, np.int64).reshape(-1)
        order = np.argsort(idx, kind="stable")
        ks = idx[order]
        uk, counts = np.unique(ks, return_counts=True)
        width = int(counts.max()) if counts.size else 0
        start = np.cumsum(counts) - counts
        seg = np.repeat(np.arange(uk.size), counts)
        pos = np.arange(ks.size) - start[seg]
        flat = torch.from_numpy(np.concatenate([idx, order, ks, uk, seg, pos])).to(device)
        n, u = idx.size, uk.size
        return SegmentPlan(flat[:n], flat[n:2 * n], flat[2 * n:3 * n], flat[3 * n:3 * n + u],
                           width, flat[3 * n + u:4 * n + u], flat[4 * n + u:])
    index = torch.as_tensor(index, device=device)
    order = torch.argsort(index, stable=True)
    ks = index[order]
    uk, counts = torch.unique_consecutive(ks, return_counts=True)
    width = int(counts.max()) if counts.numel() else 0
    start = torch.cumsum(counts, 0) - counts
    seg = torch.repeat_interleave(torch.arange(uk.numel(), device=device), counts)
    pos = torch.arange(ks.numel(), device=device) - start[seg]
    return SegmentPlan(index, order, ks, uk, width, seg, pos)


def segment_sum(index, src, n: int):
    """(n, ...) the sums of the rows of ``src`` by ``index`` (a tensor or host array of row
    numbers in [0, n), or its SegmentPlan), bit for bit the same on every call: the GPU's
    index_add adds a row's terms by atomics in any order, so its sums may differ in the last bit
    from call to call. Here the rows are sorted by index (stably), laid out one segment per row
    of a padded table, and summed along it; every scatter writes each place once.
    Differentiable."""
    out = src.new_zeros((n,) + tuple(src.shape[1:]))
    if src.shape[0] == 0:
        return out
    p = index if isinstance(index, SegmentPlan) else segment_plan(index, src.device)
    if p.width == 1:
        return out.index_add(0, p.ks, src[p.order])
    table = src.new_zeros((p.uk.numel(), p.width) + tuple(src.shape[1:]))
    table = table.index_put((p.seg, p.pos), src[p.order])
    return out.index_add(0, p.uk, table.sum(1))


def mix_parts(scalars, greek, phase, fractions, normalized: bool = True):
    """A mixture of parts, as the reference's engine mixes them
    (scatrs/src/ensemble/engine.rs:2120-2150): each part's number averages, taken from its
    outputs on the fixed path computed with ``normalized=False`` (scalars (parts, ..., 11),
    greek (parts, ..., 6, L+1) or None, phase (parts, ..., 6, n) or None), weighted by its number
    fraction (``fractions`` (parts,): the component's weight over the total times the mode's
    weight over its component's total; or (parts, ...) with the batch axes of the outputs,
    for fractions that differ between rows), and the outputs assembled from the mixed averages as
    for one part (engine.rs:2396-2497): <C_ext>, <C_sca>, <C_abs>, <C_back>, <r^2>, <r^3>
    averaged, ssa, g, the lidar ratio and the effective radius as their ratios, the phase
    matrix divided by <C_sca> and the Greek coefficients by their alpha1_0 when normalized.
    torch tensors in and out, differentiable in every input, the fractions included (with
    ``ensemble_fixed(..., normalized=False)`` for the parts, a mixture's gradient)."""
    f = fractions.to(scalars.dtype)
    f = f.reshape(f.shape + (1,) * (scalars.dim() - 1 - f.dim()))  # (parts, batch axes)
    avg = (f.unsqueeze(-1) * number_averages(scalars)).sum(0)
    ph = None if phase is None else (f[..., None, None] * phase).sum(0)
    gr = None if greek is None else (f[..., None, None] * greek).sum(0)
    return assemble(avg, gr, ph, normalized)


def number_averages(scalars):
    """The linear number averages behind the outputs of the fixed path with
    ``normalized=False`` (scalars (..., 11)): (..., 7) of <C_ext>, <C_sca>, <C_abs>, <C_back>,
    <C_sca g>, <r^2>, <r^3>, the quantities a mixture averages (engine.rs:2120-2150)."""
    import torch

    sc = scalars
    return torch.stack([sc[..., 0], sc[..., 1], sc[..., 2], sc[..., 3], sc[..., 5] * sc[..., 1],
                        sc[..., 9] / math.pi, sc[..., 10] * (3.0 / (4.0 * math.pi))], dim=-1)


def assemble(avg, greek, phase, normalized: bool = True):
    """The outputs from mixed number averages (avg (..., 7) as number_averages gives them;
    greek (..., 6, L+1) and phase (..., 6, n) mixed alike, or None), as the reference assembles
    them (engine.rs:2396-2497): ssa, g, the lidar ratio and the effective radius as ratios of
    averages, the phase matrix divided by <C_sca> and the Greek coefficients by alpha1_0 when
    normalized; depolarization 0 (spheres)."""
    import torch

    ext, sca, absb, back, scag, r2, r3 = avg.unbind(-1)
    out = torch.stack([ext, sca, absb, back, sca / ext, scag / sca, 4.0 * math.pi * ext / back,
                       torch.zeros_like(ext), r3 / r2, math.pi * r2,
                       4.0 * math.pi / 3.0 * r3], dim=-1)
    ph = phase
    if ph is not None and normalized:
        ph = ph / sca[..., None, None]
    gr = greek
    if gr is not None and normalized:
        gr = gr / gr[..., 0:1, 0:1]
    return out, gr, ph


def _parts_device(parts, wl: np.ndarray, *, order, n_radius, n_angles, allow_unvalidated,
                  want_phase, mu=None):
    """Each part (kind, p0, p1, p2, particle) at every wavelength on the fixed path with
    normalized=False, as tensors (parts, wavelengths, ...); one library call per layer count."""
    import torch

    dev = torch.device("cuda", torch.cuda.current_device())
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    nw, P = wl.size, len(parts)
    sc = torch.empty((P, nw, 11), dtype=torch.float64, device=dev)
    gr = None if order is None else torch.empty((P, nw, 6, order + 1), dtype=torch.float64,
                                                device=dev)
    n_phase = n_angles if mu is None else int(np.size(mu))
    ph = torch.empty((P, nw, 6, n_phase), dtype=torch.float64, device=dev) if want_phase \
        else None
    groups: dict[int, list[int]] = {}
    for i, part in enumerate(parts):
        groups.setdefault(part[4].layers, []).append(i)
    terms = 0
    for L, idx in groups.items():
        kind = np.repeat(np.array([parts[i][0] for i in idx], np.int32), nw)
        p0 = np.repeat(np.array([parts[i][1] for i in idx], np.float64), nw)
        p1 = np.repeat(np.array([parts[i][2] for i in idx], np.float64), nw)
        p2 = np.repeat(np.array([parts[i][3] for i in idx], np.float64), nw)
        m = np.concatenate([parts[i][4].indices(wl).reshape(L, -1).T for i in idx])  # (n, L)
        fr = None
        if L > 1:
            fr = t(np.concatenate([np.tile(parts[i][4].fractions(), (nw, 1)) for i in idx]))
        s, g, p, n = ensemble_fixed_device(
            t(kind), t(p0), t(p1), t(np.tile(wl.reshape(-1), len(idx))),
            t(m[:, 0] if L == 1 else m), fr, p2=t(p2), n_radius=n_radius, n_angles=n_angles,
            order=order, normalized=False, allow_unvalidated=allow_unvalidated,
            want_phase=want_phase, mu=mu)
        sel = torch.tensor(idx, device=dev)
        sc[sel] = s.reshape(len(idx), nw, 11)
        if gr is not None:
            gr[sel] = g.reshape(len(idx), nw, 6, order + 1)
        if ph is not None:
            ph[sel] = p.reshape(len(idx), nw, 6, n_phase)
        terms += n
    return sc, gr, ph, terms


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
    """The phase matrix at the cosines ``mu`` (the caller's, or the n_angles Gauss-Legendre
    nodes), arrays of shape grid + wavelength.shape + (len(mu),); f22 = f11 and f44 = f33 for
    spheres."""

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
    converged: np.ndarray | None = None  # the adaptive engine's (and the fixed grid's checks)
    panels: np.ndarray | None = None
    error: "EnsembleResult | None" = None  # the outputs' error estimates, same layout
    params: tuple = ()  # the parameters of d(), in the reference's names
    _d: tuple = ()

    def d(self, name: str) -> "EnsembleResult":
        """The derivatives of every output with respect to the parameter ``name`` (the
        reference's EnsembleResult.d), their error estimates in ``.error``."""
        if name not in self.params:
            raise KeyError(f"no derivative with respect to {name!r}; computed: {self.params}")
        return self._d[self.params.index(name)]


def _frozen(a: np.ndarray) -> np.ndarray:
    a.flags.writeable = False
    return a


def _result(sc, gr, ph, shape: tuple, wl: np.ndarray, n_angles: int, terms: int, mu=None,
            converged=None, panels=None):
    """EnsembleResult of outputs with rows in the C order of ``shape``; the phase matrix at the
    cosines mu, or at the n_angles Gauss-Legendre cosines."""
    s = sc.reshape(-1, 11).cpu().numpy()
    fields = [_frozen(s[:, k].reshape(shape)) for k in range(11)]
    g = None
    if gr is not None:
        a = gr.reshape(-1, 6, gr.shape[-1]).cpu().numpy()
        g = EnsembleGreek(*(_frozen(a[:, k].reshape(shape + (a.shape[-1],))) for k in range(6)))
    p = None
    if ph is not None:
        from .quadrature import gauss_legendre

        if mu is None:
            mu, _ = gauss_legendre(n_angles)
        mu = np.array(mu, np.float64).reshape(-1)
        a = ph.reshape(-1, 6, mu.size).cpu().numpy()
        p = EnsemblePhase(_frozen(mu), *(_frozen(a[:, k].reshape(shape + (mu.size,)))
                                         for k in range(6)))
    extra = {}
    if converged is not None:
        extra = dict(converged=_frozen(np.asarray(converged).reshape(shape)),
                     panels=_frozen(np.asarray(panels).reshape(shape)))
    return EnsembleResult(*fields, greek=g, phase=p, wavelength=_frozen(wl.copy()), terms=terms,
                          **extra)


def mix_units(scalars, greek, phase, unit, fractions, n_units: int, normalized: bool = True):
    """Parts' outputs with normalized=False (scalars (parts, ..., 11), greek, phase or None)
    mixed into their units (unit (parts,) indices in [0, n_units)) with the number fractions
    (parts,), as mix_parts mixes one unit: (n_units, ..., 11), greek, phase."""
    import torch

    dev = scalars.device
    u = torch.as_tensor(np.asarray(unit), device=dev)
    f = torch.as_tensor(np.asarray(fractions, np.float64), device=dev)
    f = f.reshape(f.shape + (1,) * (scalars.dim() - 1))
    add = lambda a, w: segment_sum(u, w * a, n_units)  # noqa: E731
    avg = add(number_averages(scalars), f)
    gr = None if greek is None else add(greek, f[..., None])
    ph = None if phase is None else add(phase, f[..., None])
    return assemble(avg, gr, ph, normalized)


def _part(size, particle):
    """(kind, p0, p1, p2, particle) of a distribution with scalar parameters."""
    return (size.kind, *(float(v) for v in size.params()), particle)


def _parts_of(size, particle, weight: float = 1.0):
    """The parts (kind, p0, p1, p2, particle) a distribution is integrated as and their number
    fractions times ``weight``: the distribution itself, a multimodal law's modes
    (engine.rs:143-160), or a tabulated law's segments."""
    if isinstance(size, MultiModal):
        parts, fr = [], []
        for (_, d), f in zip(size.modes, size.fractions()):
            p, q = _parts_of(d, particle, weight * f)
            parts += p
            fr += q
        return parts, fr
    if isinstance(size, Tabulated):
        segs, w = size.segments()
        return ([(SEGMENT, r0, r1, tilt, particle) for r0, r1, tilt in segs],
                list(weight * w / w.sum()))
    return [_part(size, particle)], [weight]


def _spec(size):
    """The reference's tagged tuple of a distribution with scalar parameters
    (scatrs-py/src/ensemble.rs:31-77)."""
    from .size import Gamma, LogNormal, ModifiedGamma, Monodisperse, PowerLaw

    f = float
    if isinstance(size, LogNormal):
        return ("lognormal", f(size.median_radius), f(size.width))
    if isinstance(size, Gamma):
        return ("gamma", f(size.effective_radius), f(size.effective_variance))
    if isinstance(size, ModifiedGamma):
        return ("modified_gamma", f(size.alpha), f(size.b), f(size.gamma))
    if isinstance(size, PowerLaw):
        return ("power_law", f(size.exponent), f(size.min_radius), f(size.max_radius))
    if isinstance(size, Monodisperse):
        return ("monodisperse", f(size.radius))
    if isinstance(size, Tabulated):
        return ("tabulated", np.array(size.radius), np.array(size.density))
    if isinstance(size, MultiModal):
        return ("multimodal", [(float(w), _spec(d)) for w, d in size.modes])
    raise TypeError(f"no spec for {type(size).__name__}")


def _particle_spec(particle, wl: np.ndarray):
    """The reference's tagged tuple of a particle at the wavelengths
    (scatrs-py/src/ensemble.rs:79-135)."""
    m = particle.indices(wl).reshape(particle.layers, -1)
    if particle.layers == 1:
        return ("sphere", np.ascontiguousarray(m[0]))
    return ("layered", np.asarray(particle.fractions(), np.float64), np.ascontiguousarray(m))


def _chosen(names: list, derivatives) -> list:
    """The indices of the requested parameters: all with True, none with False, or by name."""
    if derivatives is True:
        return list(range(len(names)))
    if derivatives is False or derivatives is None:
        return []
    if isinstance(derivatives, (str, bytes)):
        raise TypeError("derivatives must be a bool or a sequence of parameter names")
    out = []
    for n in derivatives:
        if n not in names:
            raise InvalidInputError(f"no parameter {n}; the parameters are {names}")
        out.append(names.index(n))
    return out


def _from_wire(raw, shape: tuple, wl: np.ndarray, mu, names: list, chosen: list, terms=0):
    """An EnsembleResult (with its errors, derivatives, converged and panels) from the wire
    format's dict, rows in the C order of ``shape``."""
    mu = None if mu is None else _frozen(np.array(mu, np.float64).reshape(-1))

    def outputs(v, g, p, **extra):
        fields = [_frozen(np.ascontiguousarray(v[k]).reshape(shape)) for k in range(11)]
        gk = None if g is None else EnsembleGreek(*(_frozen(np.ascontiguousarray(g[k]).reshape(
            shape + (g.shape[-1],))) for k in range(6)))
        pk = None if p is None else EnsemblePhase(mu, *(_frozen(np.ascontiguousarray(
            p[k]).reshape(shape + (p.shape[-1],))) for k in range(6)))
        return EnsembleResult(*fields, greek=gk, phase=pk, wavelength=_frozen(wl.copy()),
                              **extra)

    gv, ge = raw.get("greek"), raw.get("greek_errors")
    pv, pe = raw.get("phase"), raw.get("phase_errors")
    err = outputs(raw["errors"], ge, pe)
    ds = []
    for k in range(len(chosen)):
        de = outputs(raw["d_errors"][k], None if ge is None else raw["greek_d_errors"][k],
                     None if pe is None else raw["phase_d_errors"][k])
        ds.append(outputs(raw["d"][k], None if gv is None else raw["greek_d"][k],
                          None if pv is None else raw["phase_d"][k], error=de))
    return outputs(raw["values"], gv, pv, terms=terms,
                   converged=_frozen(raw["converged"].reshape(shape)),
                   panels=_frozen(raw["panels"].reshape(shape)), error=err,
                   params=tuple(names[i] for i in chosen), _d=tuple(ds))


@dataclass(frozen=True)
class Ensemble:
    """Particles of one kind over a size distribution (python/scatrs/ensemble.py:484-529).

    ``size`` may have array-valued parameters: a grid of distributions, computed at once
    (the results have the grid's axes first). A MultiModal size is integrated mode by mode, a
    Tabulated one segment by segment, and the parts mixed (mix_parts)."""

    particle: Sphere | LayeredSphere
    size: Distribution

    def __post_init__(self) -> None:
        if not isinstance(self.particle, (Sphere, LayeredSphere)):
            raise TypeError("particle must be a Sphere or a LayeredSphere, not "
                            f"{type(self.particle).__name__}")
        check(self.size)

    @property
    def grid_shape(self) -> tuple[int, ...]:
        """The shape of the grid of distributions (``()`` without array-valued parameters)."""
        if isinstance(self.size, (MultiModal, Tabulated)):
            return ()
        return np.broadcast(*self.size.params()).shape

    def compute(self, wavelength, *, greek: int | None = None, normalized: bool = True,
                mu=None, derivatives=False, rtol: float = 1e-6, max_panels: int = 4096,
                allow_unvalidated: bool = False, refine_backscatter: bool = False,
                all_orders: bool = False, n_radius: int | None = None,
                n_angles: int | None = None, phase: bool = True) -> EnsembleResult:
        """Averages over the ensemble at each wavelength, as the reference's
        Ensemble.compute (python/scatrs/ensemble.py:484-591): by the adaptive engine
        (scatrs_amd.adaptive: Gauss-Kronrod panels refined to rtol, at most max_panels per
        part; each Greek projection exact for its nodes) through the wire format, with the
        result's error, d(name), converged and panels; the phase matrix at the cosines ``mu``.
        With ``n_radius``, on the fixed discrete path instead (n_radius Gauss-Legendre nodes;
        without derivatives, the phase matrix at ``mu`` or at the n_angles Gauss-Legendre
        cosines, 128 by default). ``derivatives``: True, or the parameters' names."""
        import torch

        from . import _wire

        wl = np.asarray(wavelength, np.float64)
        if n_radius is None or derivatives:
            grid = self.grid_shape
            rt = dict(rtol=rtol, max_panels=max_panels, allow_unvalidated=allow_unvalidated,
                      refine_backscatter=refine_backscatter, n_radius=n_radius,
                      n_angles=n_angles)
            muv = np.zeros(0) if mu is None else np.asarray(mu, np.float64).reshape(-1)
            pspec = _particle_spec(self.particle, wl)
            if grid == ():
                items = [(_spec(self.size), pspec, 1.0)]
                names = _wire.ensemble_parameters(items, False)
                chosen = _chosen(names, derivatives)
                raw = _wire.ensemble_compute(
                    items, wl, False, greek, all_orders, muv, normalized, chosen,
                    rt["rtol"], rt["max_panels"], rt["allow_unvalidated"],
                    rt["refine_backscatter"], n_radius=n_radius, n_angles=n_angles)
                return _from_wire(raw, wl.shape, wl, mu, names, chosen)
            if derivatives:
                raise InvalidInputError("derivatives need a distribution with scalar parameters "
                                        "(a grid is computed without them)")
            import dataclasses

            pg = [np.broadcast_to(v, grid).reshape(-1) for v in self.size.params()]
            n_args = len(dataclasses.fields(self.size))  # in the order of params()
            specs = [_spec(type(self.size)(*(float(pg[k][i]) for k in range(n_args))))
                     for i in range(pg[0].size)]
            raw = _wire.ensemble_compute_batch(
                specs, pspec, wl, greek, all_orders, muv, normalized, rt["rtol"],
                rt["max_panels"], rt["allow_unvalidated"], rt["refine_backscatter"],
                n_radius=n_radius, n_angles=n_angles)
            return _from_wire(raw, grid + wl.shape, wl, mu, [], [])
        if n_angles is None:
            n_angles = 128
        if isinstance(self.size, (MultiModal, Tabulated)):
            parts, fr = _parts_of(self.size, self.particle)
            sc, gr, ph, terms = _parts_device(parts, wl, order=greek, n_radius=n_radius,
                                              n_angles=n_angles,
                                              allow_unvalidated=allow_unvalidated,
                                              want_phase=phase, mu=mu)
            f = torch.tensor(fr, dtype=torch.float64, device=sc.device)
            sc, gr, ph = mix_parts(sc, gr, ph, f, normalized)
            return _result(sc, gr, ph, wl.shape, wl, n_angles, terms, mu)
        p0, p1, p2 = self.size.params()
        grid = np.broadcast(p0, p1, p2).shape
        p0g, p1g, p2g = (np.broadcast_to(v, grid).reshape(-1) for v in (p0, p1, p2))
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
            t(mm[:, 0] if L == 1 else mm), fr, p2=t(np.repeat(p2g, nw)), n_radius=n_radius,
            n_angles=n_angles, order=greek, normalized=normalized,
            allow_unvalidated=allow_unvalidated, want_phase=phase, mu=mu)
        return _result(sc, gr, ph, grid + wl.shape, wl, n_angles, terms, mu)


@dataclass(frozen=True)
class Mixture:
    """An external mixture: ensembles with relative number concentrations
    (python/scatrs/ensemble.py:614-657). Cross sections and moments are averaged with the
    number fractions w_i / sum w, phase matrices with w_i times the <C_sca> of ensemble i
    (scatrs/src/ensemble/engine.rs:2120-2150); a multimodal component contributes its modes with
    its weight times their fractions.

    ``components``: (weight w_i >= 0, finite, not all 0; Ensemble of one distribution, not a
    grid) pairs."""

    components: tuple

    def __post_init__(self) -> None:
        out = []
        for i, c in enumerate(tuple(self.components)):
            try:
                w, e = c
            except (TypeError, ValueError):
                raise TypeError(f"component {i} must be a (weight, Ensemble) pair, not "
                                f"{type(c).__name__}") from None
            if not isinstance(e, Ensemble):
                raise TypeError(f"component {i} must be a (weight, Ensemble) pair, not "
                                f"({type(w).__name__}, {type(e).__name__})")
            if e.grid_shape != ():
                raise InvalidInputError(
                    f"the size of component {i} is a grid of shape {e.grid_shape}; a mixture "
                    "needs one distribution with scalar parameters per component")
            w = float(w)
            if not (np.isfinite(w) and w >= 0.0):
                raise InvalidInputError(
                    f"the weight of component {i} must be finite and >= 0, got {w}")
            out.append((w, e))
        if not out:
            raise InvalidInputError("a mixture needs a component")
        if sum(w for w, _ in out) <= 0.0:
            raise InvalidInputError("the weights of a mixture must not all be 0")
        object.__setattr__(self, "components", tuple(out))

    def compute(self, wavelength, *, greek: int | None = None, normalized: bool = True,
                mu=None, derivatives=False, rtol: float = 1e-6, max_panels: int = 4096,
                allow_unvalidated: bool = False, refine_backscatter: bool = False,
                all_orders: bool = False, n_radius: int | None = None,
                n_angles: int | None = None, phase: bool = True) -> EnsembleResult:
        """Averages over the mixture at each wavelength, as Ensemble.compute: by the adaptive
        engine through the wire format (with error, d(name) in the mixture's parameter names
        c{i}.*, converged, panels), or on the fixed discrete path with ``n_radius``."""
        import torch

        from . import _wire

        wl = np.asarray(wavelength, np.float64)
        if n_radius is None or derivatives:
            items = [(_spec(e.size), _particle_spec(e.particle, wl), w)
                     for w, e in self.components]
            names = _wire.ensemble_parameters(items, True)
            chosen = _chosen(names, derivatives)
            muv = np.zeros(0) if mu is None else np.asarray(mu, np.float64).reshape(-1)
            raw = _wire.ensemble_compute(items, wl, True, greek, all_orders, muv, normalized,
                                         chosen, rtol, max_panels, allow_unvalidated,
                                         refine_backscatter, n_radius=n_radius,
                                         n_angles=n_angles)
            return _from_wire(raw, wl.shape, wl, mu, names, chosen)
        total = sum(w for w, _ in self.components)
        parts, fr = [], []
        for w, e in self.components:
            if w == 0.0:
                continue
            p, f = _parts_of(e.size, e.particle, w / total)
            parts += p
            fr += f
        if n_angles is None:
            n_angles = 128
        sc, gr, ph, terms = _parts_device(parts, wl, order=greek, n_radius=n_radius,
                                          n_angles=n_angles, allow_unvalidated=allow_unvalidated,
                                          want_phase=phase, mu=mu)
        f = torch.tensor(fr, dtype=torch.float64, device=sc.device)
        sc, gr, ph = mix_parts(sc, gr, ph, f, normalized)
        return _result(sc, gr, ph, wl.shape, wl, n_angles, terms, mu)
