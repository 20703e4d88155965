"""The reference's ``_scatrs`` ensemble entry points over the fixed discrete path.

``ensemble_parameters``, ``ensemble_compute`` and ``ensemble_compute_batch`` take what the
reference's bindings take (scatrs-py/src/ensemble.rs:172-504, python/scatrs/_scatrs/
__init__.pyi:141-168): distributions and particles as tagged tuples (ensemble.rs:31-135),
components as (distribution spec, particle spec, weight), the options in the same order. They
return the same dict of arrays, one row per wavelength (``ensemble_compute_batch``: per
(distribution, wavelength), distribution-major):

* ``values``, ``errors`` (11, n) in SCALARS order; ``d``, ``d_errors`` (P, 11, n);
* ``greek``, ``greek_errors`` (6, n, L+1) and ``greek_d``, ``greek_d_errors`` (P, 6, n, L+1)
  with a Greek order;
* ``phase``, ``phase_errors`` (6, n, n_mu) and ``phase_d``, ``phase_d_errors`` (P, 6, n, n_mu)
  with cosines mu (F11, F12, F22, F33, F34, F44);
* ``converged`` bool (n) and ``panels`` uint64 (n).

By default they run the adaptive engine (scatrs_amd.adaptive), which takes rtol,
max_panels and refine_backscatter as the reference does. Given ``n_radius``, they run the fixed
path instead. There each part (a distribution, a multimodal law's mode, a tabulated law's
segment) is one ``n_radius``-point Gauss-Legendre panel in ln r, and the following stand in for
the reference's adaptive integration (engine.rs:307-943):

* the error estimate of every output is its difference from the same computation on
  ``n_radius // 2`` nodes (the estimate of the coarser rule, as the reference's Kronrod minus
  Gauss difference is; the adaptive engine's are the reference's bounds, _error_bounds);
* ``converged`` applies the reference's refinement targets (engine.rs:620-690) to those
  differences part by part: C_ext, C_sca and C_abs (against C_ext's rtol), C_sca g against
  C_sca's, and with ``refine_backscatter`` C_back against the larger of C_back and C_sca;
* ``panels`` is the number of parts (one panel each); ``max_panels`` is checked, not used.

Derivatives are the reference's forward-mode columns (the parameters named by
``ensemble_parameters``, each result differentiated in its own wavelength's parameters), here
from one reverse pass: every output component of a wavelength gets a replica of the whole
computation with a one-hot cotangent, so one call of the library's reverse-mode kernel returns
every (output, parameter) pair. The size range is held fixed, as the reference differentiates
(``fixed_range``).
"""

from __future__ import annotations

import numpy as np

from .ensemble import segment_plan, segment_sum
from .errors import InvalidInputError
from .size import (SEGMENT, Gamma, LogNormal, ModifiedGamma, Monodisperse, MultiModal,
                   PowerLaw, Tabulated)

__all__ = ["MAX_GREEK_ORDER", "MAX_MIE_VALUES", "ensemble_compute", "ensemble_compute_batch",
           "ensemble_parameters"]

MAX_MIE_VALUES = 1 << 27  # scatrs-py/src/mie.rs:30
MAX_GREEK_ORDER = 2 * (1 << 20)  # scatrs/src/mie/greek.rs:344, mie/mod.rs:101
_N = 11  # scalar outputs


def _distribution(spec):
    """A distribution from its tagged tuple (scatrs-py/src/ensemble.rs:31-77)."""
    if not isinstance(spec, tuple) or not spec:
        raise TypeError("a distribution spec must be a tuple")
    kind = spec[0]
    f = lambda i: float(spec[i])  # noqa: E731
    if kind == "lognormal":
        return LogNormal(f(1), f(2))
    if kind == "gamma":
        return Gamma(f(1), f(2))
    if kind == "modified_gamma":
        return ModifiedGamma(f(1), f(2), f(3))
    if kind == "power_law":
        return PowerLaw(f(1), f(2), f(3))
    if kind == "monodisperse":
        return Monodisperse(f(1))
    if kind == "tabulated":
        return Tabulated(tuple(np.asarray(spec[1], np.float64).reshape(-1)),
                         tuple(np.asarray(spec[2], np.float64).reshape(-1)))
    if kind == "multimodal":
        return MultiModal(tuple((float(w), _distribution(s)) for w, s in spec[1]))
    raise TypeError(f"unknown distribution kind {kind!r}")


def _particle(spec):
    """(m (L, n) complex, radius fractions (L-1,)) from a particle's tagged tuple
    (scatrs-py/src/ensemble.rs:79-135)."""
    if not isinstance(spec, tuple) or not spec:
        raise TypeError("a particle spec must be a tuple")
    kind = spec[0]
    if kind == "sphere":
        if not isinstance(spec[1], np.ndarray):
            raise TypeError("a sphere's refractive indices must be an array")
        return np.asarray(spec[1], np.complex128).reshape(1, -1), np.zeros(0)
    if kind == "layered":
        if not isinstance(spec[1], np.ndarray):
            raise TypeError("a layered sphere's radius fractions must be an array")
        if not isinstance(spec[2], np.ndarray):
            raise TypeError("a layered sphere's refractive indices must be an array")
        m = np.asarray(spec[2], np.complex128)
        if m.ndim != 2:
            raise TypeError("a layered sphere's refractive indices must have shape (layers, "
                            f"wavelengths), got {m.ndim} axes")
        fr = np.asarray(spec[1], np.float64).reshape(-1)
        if fr.size != m.shape[0] - 1:
            raise InvalidInputError(f"a layered sphere of {m.shape[0]} layers needs "
                                    f"{m.shape[0] - 1} radius fractions, got {fr.size}")
        return m, fr
    if kind == "spheroid":
        raise InvalidInputError("spheroids are not on the ensemble path of this library "
                                "(spheres and layered spheres)")
    raise TypeError(f"unknown particle kind {kind!r}")


class _Component:
    """One component: its distribution, m (L, n), radius fractions and weight."""

    def __init__(self, dist, m, fr, weight):
        self.dist, self.m, self.fr, self.weight = dist, m, fr, float(weight)

    @property
    def layers(self) -> int:
        return self.m.shape[0]


def _components(items):
    return [_Component(_distribution(s), *_particle(p), w) for s, p, w in items]


def _check_distribution(dist) -> None:
    """The reference's parameter checks of a distribution and its modes
    (distribution.rs:99-185), before any device work."""
    from .adaptive import check_parameters

    if isinstance(dist, MultiModal):
        for j, (_, d) in enumerate(dist.modes):
            try:
                _check_distribution(d)
            except InvalidInputError as e:
                raise InvalidInputError(f"mode {j}: {e}") from None
    elif not isinstance(dist, Tabulated):
        check_parameters(dist.kind, np.array(dist.params(), np.float64))


def _validate(comps, n: int) -> None:
    """The reference's checks of the components (scatrs/src/ensemble/mod.rs:745-780): valid
    distributions, one refractive index per wavelength, finite weights >= 0 with a positive
    sum."""
    total = 0.0
    for i, c in enumerate(comps):
        _check_distribution(c.dist)
        if c.m.shape[1] != n:
            raise InvalidInputError(
                f"component {i} has {c.m.shape[1]} refractive indices for {n} wavelengths")
        if not (np.isfinite(c.weight) and c.weight >= 0.0):
            raise InvalidInputError(
                f"the weight of component {i} must be finite and >= 0, got {c.weight}")
        total += c.weight
    if not (np.isfinite(total) and total > 0.0):
        raise InvalidInputError(
            f"the mixture weights must have a finite, positive sum, got {total}")


def _component_parameters(dist, layers: int, weight: bool):
    """(name, (mode, what, index)) of a component's parameters, in the reference's order
    (scatrs/src/ensemble/mod.rs:252-312): the distribution's (each mode's, then the mode's
    weight), the particle's (Re m, Im m of each layer, then the radius fractions), then the
    component's weight in a mixture."""
    out = []
    if isinstance(dist, MultiModal):
        for j, (_, d) in enumerate(dist.modes):
            out += [(f"mode{j}.{n}", (j, "dist", s)) for s, n in enumerate(d.names)]
            out.append((f"mode{j}.weight", (j, "weight", 0)))
    else:
        out += [(n, (None, "dist", s)) for s, n in enumerate(dist.names)]
    if layers == 1:
        out += [("m_real", (None, "m_real", 0)), ("m_imag", (None, "m_imag", 0))]
    else:
        for l in range(layers):
            out += [(f"layer{l}.m_real", (None, "m_real", l)),
                    (f"layer{l}.m_imag", (None, "m_imag", l))]
        out += [(f"layer{l}.radius_fraction", (None, "fraction", l)) for l in range(layers - 1)]
    if weight:
        out.append(("weight", (None, "weight", 0)))
    return out


def _parameters(comps, mixture: bool):
    """The names and keys (component, mode, what, index) of every parameter
    (scatrs-py/src/ensemble.rs:154-170; names of scatrs/src/ensemble/mod.rs:225-249)."""
    if not mixture and len(comps) != 1:
        raise InvalidInputError(
            f"a single distribution needs exactly one component, got {len(comps)}")
    names, keys = [], []
    for i, c in enumerate(comps):
        for n, k in _component_parameters(c.dist, c.layers, mixture):
            names.append(f"c{i}.{n}" if mixture else n)
            keys.append((i,) + k)
    return names, keys


def ensemble_parameters(items, mixture: bool) -> list[str]:
    """The names of the parameters of the components (a single distribution unless
    ``mixture``), in the order ``ensemble_compute`` indexes them
    (scatrs-py/src/ensemble.rs:172-187)."""
    comps = _components(items)
    if not comps:
        return []
    return _parameters(comps, mixture)[0]


def _order(greek, all_orders: bool):
    """The Greek order, "all", or None (scatrs-py/src/ensemble.rs:371-381)."""
    if all_orders:
        return "all"
    if greek is None:
        return None
    if greek < 0 or greek > MAX_GREEK_ORDER:
        raise InvalidInputError(f"the Greek order must be at most {MAX_GREEK_ORDER}, got {greek}")
    return int(greek)


def _check_options(rtol, max_panels, n, n_mu, order, p):
    """The reference's option checks (scatrs/src/ensemble/mod.rs:689-703) and its bound on
    the outputs (scatrs-py/src/ensemble.rs:382-393; every order: known after pass 1)."""
    if order == "all":
        order = None
    if not (1e-13 <= rtol <= 0.1):
        raise InvalidInputError(f"rtol must be in [1e-13, 0.1], got {rtol}")
    if max_panels < 1:
        raise InvalidInputError("max_panels must be at least 1")
    per = _N + 4 * n_mu + (0 if order is None else 6 * (order + 1))
    if n * per * 2 * (1 + p) > 2 * MAX_MIE_VALUES:
        raise InvalidInputError("the ensemble outputs would exceed 2 * MAX_MIE_VALUES = "
                                f"{2 * MAX_MIE_VALUES} real values")


class _Leaves:
    """The parameters of the units' components as tensors with one row per (replica,
    wavelength): dist[(u, c, mode)] (B, 3), m_real/m_imag[(u, c)] (B, L), fraction[(u, c)]
    (B, L-1), weight[(u, c)] (B,), mode_weight[(u, c)] (B, modes)."""

    def __init__(self, units, n: int, R: int, dev, grad: bool):
        import torch

        B = R * n
        self.n = n
        col = np.tile(np.arange(n), R)
        self.dist, self.m_real, self.m_imag, self.fraction = {}, {}, {}, {}
        self.weight, self.mode_weight = {}, {}
        specs = []  # (table, key, host array): all go to the device in one transfer

        def t(table, key, a):
            specs.append((table, key, np.ascontiguousarray(a, dtype=np.float64)))

        for u, comps in enumerate(units):
            for c, comp in enumerate(comps):
                mb = comp.m[:, col].T  # (B, L)
                t(self.m_real, (u, c), mb.real)
                t(self.m_imag, (u, c), mb.imag)
                t(self.fraction, (u, c), np.tile(comp.fr, (B, 1)))
                t(self.weight, (u, c), np.full(B, comp.weight))
                d = comp.dist
                if isinstance(d, MultiModal):
                    t(self.mode_weight, (u, c), np.tile([w for w, _ in d.modes], (B, 1)))
                    for j, (_, dj) in enumerate(d.modes):
                        if not isinstance(dj, Tabulated):
                            t(self.dist, (u, c, j),
                              np.tile(np.array(dj.params(), np.float64), (B, 1)))
                elif not isinstance(d, Tabulated):
                    t(self.dist, (u, c, None),
                      np.tile(np.array(d.params(), np.float64), (B, 1)))
        flat = np.concatenate([a.reshape(-1) for _, _, a in specs]) if specs else np.zeros(0)
        # Leaves of reverse mode: views of one tensor that requires grad, each a graph node of
        # its own (its gradient is the slice's).
        base = torch.from_numpy(flat).to(dev).requires_grad_(grad)
        off = 0
        for table, key, a in specs:
            table[key] = base[off:off + a.size].view(a.shape)
            off += a.size

    def tensors(self):
        return [*self.dist.values(), *self.m_real.values(), *self.m_imag.values(),
                *self.fraction.values(), *self.weight.values(), *self.mode_weight.values()]

    def _slot(self, u: int, key):
        """(table, its key, column or None for a vector) of the parameter ``key`` =
        (component, mode, what, index) of unit u."""
        c, j, what, i = key
        if what == "dist":
            return self.dist, (u, c, j), i
        if what in ("m_real", "m_imag", "fraction"):
            return getattr(self, what), (u, c), i
        if j is None:
            return self.weight, (u, c), None
        return self.mode_weight, (u, c), j

    def column(self, u: int, key, grads: dict):
        """The gradient column (B,) of the parameter ``key`` = (component, mode, what, index) of
        unit u, from ``grads`` (id(tensor) -> its gradient or None)."""
        import torch

        table, k, col = self._slot(u, key)
        leaf = table[k]
        g = grads.get(id(leaf))
        if g is None:
            return torch.zeros(leaf.shape[0], dtype=torch.float64, device=leaf.device)
        return g if col is None else g[:, col]

    def seed(self, u: int, keys):
        """Forward mode: replica r of unit u's parameter keys[r] gets tangent 1, every other
        row and parameter 0 (inside a torch.autograd.forward_ad.dual_level)."""
        import torch
        from torch.autograd import forward_ad

        tangents = {}
        for r, key in enumerate(keys):
            table, k, col = self._slot(u, key)
            tg = tangents.setdefault((id(table), k), (table, k, torch.zeros_like(table[k])))[2]
            rows = slice(r * self.n, (r + 1) * self.n)
            if col is None:
                tg[rows] = 1.0
            else:
                tg[rows, col] = 1.0
        for table, k, tg in tangents.values():
            table[k] = forward_ad.make_dual(table[k], tg)


def _parts(units, leaves: _Leaves):
    """Every part (unit, kind, p (B, 3), m_real, m_imag, fractions, number fraction (B,),
    layers, (component, mode)) of the units: a component's distribution, a multimodal law's modes
    (engine.rs:143-160) or a tabulated law's segments, with the number fractions of the
    reference's mixing (engine.rs:2131-2133): the component's weight over the unit's total
    times the mode's weight over its component's total, times a segment's share of its
    tabulated law's number."""
    import torch

    out = []
    for u, comps in enumerate(units):
        total = sum(leaves.weight[u, c] for c in range(len(comps)))
        for c, comp in enumerate(comps):
            fc = leaves.weight[u, c] / total
            m_re, m_im, fr = leaves.m_real[u, c], leaves.m_imag[u, c], leaves.fraction[u, c]
            modes = [(None, comp.dist, fc)]
            if isinstance(comp.dist, MultiModal):
                v = leaves.mode_weight[u, c]
                share = v / v.sum(1, keepdim=True)
                modes = [(j, d, fc * share[:, j]) for j, (_, d) in enumerate(comp.dist.modes)]
            for j, d, f in modes:
                if isinstance(d, Tabulated):
                    segs, w = d.segments()
                    w = w / w.sum()
                    for (r0, r1, tilt), ws in zip(segs, w):
                        p = torch.tensor([r0, r1, tilt], dtype=torch.float64,
                                         device=f.device).expand(f.shape[0], 3)
                        out.append((u, SEGMENT, p, m_re, m_im, fr, f * ws, comp.layers,
                                    (c, j)))
                else:
                    out.append((u, d.kind, leaves.dist[u, c, j], m_re, m_im, fr, f,
                                comp.layers, (c, j)))
    return out


def _evaluate(units, wl: np.ndarray, *, R: int, grad: bool, order, mu, normalized: bool,
              n_radius: int, n_angles: int, allow_unvalidated: bool):
    """The outputs of every unit at R replicas of the wavelengths on the fixed path: (leaves,
    scalars (U, B, 11), greek (U, B, 6, L+1) or None, phase (U, B, 6, n_mu) or None, the parts'
    scalars (parts, B, 11), each part's unit)."""
    import torch

    from .ensemble import assemble, ensemble_fixed, number_averages

    dev = torch.device("cuda", torch.cuda.current_device())
    n = wl.size
    B = R * n
    leaves = _Leaves(units, n, R, dev, grad)
    parts = _parts(units, leaves)
    wl_rows = torch.tensor(np.tile(wl, R), dtype=torch.float64, device=dev)
    phase = mu is not None
    by_layers: dict[int, list[int]] = {}
    for i, part in enumerate(parts):
        by_layers.setdefault(part[7], []).append(i)
    sc_of, gr_of, ph_of = [None] * len(parts), [None] * len(parts), [None] * len(parts)
    for L, idx in by_layers.items():
        k = len(idx)
        kind = torch.tensor(np.repeat([parts[i][1] for i in idx], B), dtype=torch.int32,
                            device=dev)
        p = torch.cat([parts[i][2] for i in idx])
        m_re = torch.cat([parts[i][3] for i in idx])
        m_im = torch.cat([parts[i][4] for i in idx])
        fr = torch.cat([parts[i][5] for i in idx]) if L > 1 else None
        if L == 1:
            m_re, m_im = m_re[:, 0], m_im[:, 0]
        sc, gr, ph = ensemble_fixed(
            kind, p[:, 0], p[:, 1], wl_rows.repeat(k), m_re, m_im, fr, p2=p[:, 2],
            n_radius=n_radius, n_angles=n_angles, order=order, normalized=False,
            fixed_range=True, allow_unvalidated=allow_unvalidated, mu=mu, phase=phase)
        for q, i in enumerate(idx):
            sc_of[i] = sc[q * B:(q + 1) * B]
            gr_of[i] = None if gr is None else gr[q * B:(q + 1) * B]
            ph_of[i] = None if ph is None else ph[q * B:(q + 1) * B]
    U = len(units)
    unit = segment_plan(np.array([part[0] for part in parts], np.int64), dev)
    f = torch.stack([part[6] for part in parts])  # (parts, B)
    sc_parts = torch.stack(sc_of)
    avg = segment_sum(unit, f[..., None] * number_averages(sc_parts), U)
    gr = None
    if order is not None:
        g = torch.stack(gr_of)
        gr = segment_sum(unit, f[..., None, None] * g, U)
    ph = None
    if phase:
        q = torch.stack(ph_of)
        ph = segment_sum(unit, f[..., None, None] * q, U)
    sc, gr, ph = assemble(avg, gr, ph, normalized)
    return leaves, sc, gr, ph, sc_parts.detach(), [part[0] for part in parts]


def _flat(sc, gr, ph):
    """(rows, K): the scalars, then the Greek coefficients, then the phase matrix."""
    import torch

    cols = [sc]
    if gr is not None:
        cols.append(gr.reshape(gr.shape[0], -1))
    if ph is not None:
        cols.append(ph.reshape(ph.shape[0], -1))
    return torch.cat(cols, 1)


def _converged(sc_parts, sc_half, part_unit, U: int, rtol: float, refine_backscatter: bool):
    """(U, rows) bool: every part of a unit meets the reference's refinement targets
    (engine.rs:620-690) with the difference of the two rules as its error."""
    import torch

    ext, sca, absb, back = (sc_parts[..., k] for k in range(4))
    scag = sc_parts[..., 5] * sca
    d = (sc_parts - sc_half).abs()
    d_scag = (scag - sc_half[..., 5] * sc_half[..., 1]).abs()
    ok = (d[..., 0] <= rtol * ext.abs()) & (d[..., 1] <= rtol * sca.abs())
    ok &= (d[..., 2] <= rtol * ext.abs()) & (d_scag <= rtol * sca.abs())
    if refine_backscatter:
        ok &= d[..., 3] <= rtol * torch.maximum(back.abs(), sca.abs())
    out = torch.ones((U, sc_parts.shape[1]), dtype=torch.bool, device=sc_parts.device)
    for p, u in enumerate(part_unit):
        out[u] &= ok[p]
    return out


def _run(units, wl, *, order, mu, normalized, keys, rtol, refine_backscatter, n_radius, n_angles,
         allow_unvalidated):
    """Values, derivatives (when ``keys``, for unit 0) and their two-rule errors; converged and
    panels."""
    import torch

    n = wl.size
    W = 0 if order is None else order + 1
    n_mu = 0 if mu is None else mu.size
    K = _N + 6 * W + 6 * n_mu
    R = K if keys else 1

    def one(nr):
        with torch.enable_grad() if keys else torch.no_grad():
            leaves, sc, gr, ph, sc_parts, part_unit = _evaluate(
                units, wl, R=R, grad=bool(keys), order=order, mu=mu, normalized=normalized,
                n_radius=nr, n_angles=n_angles, allow_unvalidated=allow_unvalidated)
            Y = [_flat(sc[u], None if gr is None else gr[u], None if ph is None else ph[u])
                 for u in range(len(units))]  # (B, K) per unit
            D = None
            if keys:
                ts = leaves.tensors()
                cot = torch.zeros((R, n, K), dtype=torch.float64, device=Y[0].device)
                cot[torch.arange(K), :, torch.arange(K)] = 1.0
                gs = torch.autograd.grad(Y[0], ts, cot.reshape(R * n, K), allow_unused=True)
                grads = {id(t): g for t, g in zip(ts, gs)}
                D = torch.stack([leaves.column(0, k, grads).reshape(R, n) for k in keys])
        V = torch.stack([y.detach()[:n] for y in Y])  # (U, n, K), replica 0
        return V, D, sc_parts[:, :n], part_unit

    V, D, sc_parts, part_unit = one(n_radius)
    Vh, Dh, sc_half, _ = one(max(1, n_radius // 2))
    conv = _converged(sc_parts, sc_half, part_unit, len(units), rtol, refine_backscatter)
    if n_radius < 2:
        conv[:] = False
    counts = np.bincount(np.asarray(part_unit), minlength=len(units))
    return V, (V - Vh).abs(), D, None if D is None else (D - Dh).abs(), conv, counts


def _dict(V, E, D, DE, conv, counts, n_per_unit: int, W: int, n_mu: int):
    """The reference's output dict (scatrs-py/src/ensemble.rs:456-478) from values and
    errors (U, n, K), derivatives and errors (P, K, n) or None."""
    U = V.shape[0]
    n = U * n_per_unit
    P = 0 if D is None else D.shape[0]
    v, e = (a.reshape(n, -1).cpu().numpy() for a in (V, E))
    if D is None:
        d = de = np.zeros((0, v.shape[1], n))
    else:
        d, de = D.cpu().numpy(), DE.cpu().numpy()
    out = {"values": np.ascontiguousarray(v[:, :_N].T), "errors": np.ascontiguousarray(e[:, :_N].T),
           "d": np.ascontiguousarray(d[:, :_N]), "d_errors": np.ascontiguousarray(de[:, :_N])}
    if W:
        s = slice(_N, _N + 6 * W)
        out["greek"], out["greek_errors"] = (
            np.ascontiguousarray(a[:, s].reshape(n, 6, W).transpose(1, 0, 2)) for a in (v, e))
        out["greek_d"], out["greek_d_errors"] = (
            np.ascontiguousarray(a[:, s].reshape(P, 6, W, n).transpose(0, 1, 3, 2))
            for a in (d, de))
    if n_mu:
        s = slice(_N + 6 * W, _N + 6 * W + 6 * n_mu)
        out["phase"], out["phase_errors"] = (
            np.ascontiguousarray(a[:, s].reshape(n, 6, n_mu).transpose(1, 0, 2)) for a in (v, e))
        out["phase_d"], out["phase_d_errors"] = (
            np.ascontiguousarray(a[:, s].reshape(P, 6, n_mu, n).transpose(0, 1, 3, 2))
            for a in (d, de))
    out["converged"] = conv.reshape(-1).cpu().numpy().astype(bool)
    out["panels"] = np.repeat(counts.astype(np.uint64), n_per_unit)
    return out


def _refined_derivatives(keys, cj, layers: int):
    """The derivatives pass 1 refines for the part (component, mode) ``cj`` (engine.rs:158-178):
    its distribution's requested parameters (indices into its p) and its particle's requested
    directions (2 l, 2 l + 1: Re m, Im m of layer l; 2 L + j: radius fraction j). Weights are
    only integrated."""
    c, j = cj
    theta, dirs = [], []
    for (kc, kj, what, i) in keys or ():
        if kc != c:
            continue
        if what == "dist" and kj == j:
            theta.append(i)
        elif what == "m_real":
            dirs.append(2 * i)
        elif what == "m_imag":
            dirs.append(2 * i + 1)
        elif what == "fraction":
            dirs.append(2 * layers + i)
    return tuple(theta), tuple(sorted(dirs))


def _run_adaptive(units, wl, *, order, mu, normalized, keys, rtol, max_panels,
                  refine_backscatter, n_angles, allow_unvalidated, shared: bool = False,
                  forward: bool | None = None):
    """Values and errors (U, n, K), derivatives and their errors (P, K, n) for unit 0 (None
    without keys), converged (U, n) and panels per (unit, wavelength) by the adaptive engine
    (scatrs_amd.adaptive): pass 1 on the parameters' values, pass 2 differentiable
    (adaptive.pass2_torch) by both rules.

    Derivatives come in forward mode (torch.autograd.forward_ad) where unit 0 has no
    monodisperse part: one replica per requested parameter, each seeded with its own one-hot
    tangent, and one dual pass 2 gives every column (the node calls' tangents come from
    scatrs_ensemble_nodes_tangent). Otherwise (or with ``forward=False``) in reverse mode: K
    replicas with one-hot cotangents, as in _run."""
    import contextlib

    import torch
    from torch.autograd import forward_ad

    from . import adaptive
    from .ensemble import assemble
    from .size import MONODISPERSE

    dev = torch.device("cuda", torch.cuda.current_device())
    n = wl.size
    W = 0 if order is None else order + 1
    n_mu = 0 if mu is None else mu.size
    K = _N + 6 * W + 6 * n_mu
    U = len(units)
    with torch.no_grad():
        probe = _parts(units, _Leaves(units, n, 1, dev, False))
    if forward is None:
        forward = not any(part[0] == 0 and int(part[1]) == MONODISPERSE for part in probe)
    fwd = bool(keys) and forward
    R = (len(keys) if fwd else K) if keys else 1
    B = R * n
    rev = bool(keys) and not fwd
    with torch.enable_grad() if rev else torch.no_grad():
        leaves = _Leaves(units, n, R, dev, rev)
        parts = _parts(units, leaves)
        tasks = []
        for (u, kind, p, m_re, m_im, fr, _, L, cj) in parts:
            p0 = p.detach().cpu().numpy()
            m0 = (m_re.detach() + 1j * m_im.detach()).cpu().numpy()
            f0 = fr.detach().cpu().numpy()
            theta, dirs = _refined_derivatives(keys if u == 0 else (), cj, L)
            assert p.shape[0] == B  # rows b = r n + i: wavelength i's replicas are [i::n]
            for i in range(n):
                tasks.append(adaptive.Task(int(kind), p0[i].copy(), float(wl[i]), m0[i].copy(),
                                           f0[i].copy(), theta=theta, dirs=dirs))
        if shared and not keys:
            # one particle for every unit: at each wavelength all parts share their panels and
            # nodes (the reference's batch, engine.rs:257-296); task k is wavelength k % n
            from . import shared as sh

            groups = [list(range(i, len(tasks), n)) for i in range(n)]
            sh.pass1_groups(tasks, groups, rtol, max_panels, refine_backscatter,
                            allow_unvalidated)
        else:
            adaptive.pass1(tasks, rtol, max_panels, refine_backscatter, allow_unvalidated)
        P = len(parts)
        out = []
        rows = []
        with forward_ad.dual_level() if fwd else contextlib.nullcontext():
            if fwd:
                leaves.seed(0, keys)
                parts = _parts(units, leaves)
            params = [(p[i::n], m_re[i::n], m_im[i::n], fr[i::n] if L > 1 else None)
                      for (_, _, p, m_re, m_im, fr, _, L, _) in parts for i in range(n)]
            unit = segment_plan(np.array([part[0] for part in parts], np.int64), dev)
            f = torch.stack([part[6] for part in parts])  # (parts, B)
            kw = dict(order=order, n_angles=n_angles, mu=mu, want_phase=mu is not None,
                      allow_unvalidated=allow_unvalidated)
            if rev:  # one differentiable pass per rule (their reverse passes stay apart)
                passes = [adaptive.pass2_torch(tasks, params, R=R, idx=idx, **kw)
                          for idx in (np.arange(15), adaptive._GAUSS)]
            elif shared and not keys:  # every shared node solved once; both rules from it
                passes = sh.pass2_groups(tasks, **kw)
            elif (not keys and order is None and mu is None
                  and all(q.kind != MONODISPERSE and q.a.size for q in tasks)):
                # no Greek, phase or derivatives: pass 1's panel sums are pass 2's
                passes = adaptive.pass2_from_panels(tasks)
            else:  # both rules in one call per angle rule: the tasks twice (the same objects:
                # pass 2 only reads them), Gauss weights second
                passes = _split_rules(adaptive.pass2_torch(
                    tasks + tasks, params + params, R=R,
                    idx=np.arange(15), gauss_tasks=np.repeat([False, True], len(tasks)),
                    derivatives=fwd, **kw), len(tasks), R)
            for avg, gr, ph, _, _, rw in passes:
                rows.append(rw)
                # (P n, R, ...) -> (P, B, ...), b = r n + i
                per = lambda a: None if a is None else a.reshape(  # noqa: E731
                    (P, n, R) + a.shape[2:]).transpose(1, 2).reshape((P, B) + a.shape[2:])
                avg, gr, ph = per(avg), per(gr), per(ph)
                mix = lambda a, w: segment_sum(unit, w * a, U)  # noqa: E731
                sc_u, gr_u, ph_u = assemble(mix(avg, f[..., None]),
                                            None if gr is None else mix(gr, f[..., None, None]),
                                            None if ph is None else mix(ph, f[..., None, None]),
                                            normalized)
                Y = [_flat(sc_u[u], None if gr_u is None else gr_u[u],
                           None if ph_u is None else ph_u[u]) for u in range(U)]
                D = None
                if rev:
                    ts = leaves.tensors()
                    cot = torch.zeros((R, n, K), dtype=torch.float64, device=dev)
                    cot[torch.arange(K), :, torch.arange(K)] = 1.0
                    # both rules share the parameters' graph (the mixing fractions): keep it
                    gs = torch.autograd.grad(Y[0], ts, cot.reshape(B, K), allow_unused=True,
                                             retain_graph=True)
                    grads = {id(t): g for t, g in zip(ts, gs)}
                    D = torch.stack([leaves.column(0, k, grads).reshape(R, n) for k in keys])
                elif fwd:
                    tg = forward_ad.unpack_dual(Y[0]).tangent
                    D = (torch.zeros((R, n, K), dtype=torch.float64, device=dev) if tg is None
                         else tg.reshape(R, n, K).permute(0, 2, 1).contiguous())
                out.append((torch.stack([y.detach()[:n] for y in Y]), D))
                if len(out) == 1:
                    kept = (avg.detach(), None if gr is None else gr.detach(),
                            None if ph is None else ph.detach(), f.detach())
    (V, D), (Vg, Dg) = out
    E = _error_bounds(rows, kept, unit, U, P, n, R, normalized)
    conv = np.ones((U, n), bool)
    panels = np.zeros((U, n), np.int64)
    for k, q in enumerate(tasks):
        u, i = parts[k // n][0], k % n
        conv[u, i] &= q.converged
        panels[u, i] += q.n_panels
    DE = None if D is None else (D - Dg).abs()
    return V, E, D, DE, torch.from_numpy(conv), panels


def _split_rules(res, T: int, R: int):
    """The Kronrod and Gauss halves of pass2_torch over a doubled task list (tasks, then their
    Gauss copies): outputs by task, and the rows' sums aligned by place. The rows are [the
    originals' panel rows, the copies' in the same order, the originals' monodisperse rows,
    the copies'], each x R; a copy's key is its original's."""
    import torch

    avg, gr, ph, terms, na, rw = res
    half = lambda a, k: None if a is None else a[k * T:(k + 1) * T]  # noqa: E731
    nr, nm = rw["n_rows"] // 2, rw["n_mono"] // 2
    dev = rw["key"].device
    out = []
    for k in range(2):
        sel = torch.cat([torch.arange((k * nr) * R, (k * nr + nr) * R, device=dev),
                         torch.arange((2 * nr + k * nm) * R, (2 * nr + k * nm + nm) * R,
                                      device=dev)])
        key = rw["key"][sel]
        if k:
            key = torch.where(key >= T * R, key - T * R, key)
        d = {"key": key, "lin": rw["lin"][sel],
             "gr": None if rw["gr"] is None else rw["gr"][sel],
             "ph": None if rw["ph"] is None else rw["ph"][sel],
             "z": rw["z"][k * T * R:(k + 1) * T * R]}
        out.append((half(avg, k), half(gr, k), half(ph, k), terms, na, d))
    return out


def _error_bounds(rows, kept, unit, U, P, n, R, normalized):
    """The values' error bounds as the reference forms them (lin.rs:7-68): each linear sum's
    error is the sum over pass-2 rows of |Kronrod - Gauss| (engine.rs:1697-1763), over the
    moment integration's Z; a mixture adds its parts' errors times |fraction|; a ratio a/b
    takes (e_a + |a/b| e_b)/|b| (ssa, g, the lidar ratio, the normalized Greek coefficients and
    phase matrix). (U, n, K), replica 0."""
    import math

    import torch

    (rk, rg), (avg, gr, ph, f) = rows, kept
    T = P * n
    dev = avg.device
    kplan = segment_plan(rk["key"], dev)

    def task_err(a, b):
        if a is None:
            return None
        d = (a - b).abs()
        return segment_sum(kplan, d, T * R)

    z = rk["z"]
    e_lin = task_err(rk["lin"], rg["lin"]) / z[:, None]
    e_gr = None if rk["gr"] is None else task_err(rk["gr"], rg["gr"]) / z[:, None, None]
    e_ph = None if rk["ph"] is None else task_err(rk["ph"], rg["ph"]) / z[:, None, None]
    # (T R, ...) -> (P, n, ...) at replica 0
    first = lambda a: None if a is None else a.reshape((P, n, R) + a.shape[1:])[:, :, 0]  # noqa
    e_lin, e_gr, e_ph = first(e_lin), first(e_gr), first(e_ph)
    e_lin[..., 5:] = 0.0  # <r^2>, <r^3>: the moment integration's, errors below rtol/10
    vals = lambda a: None if a is None else a.reshape((P, R, n) + a.shape[2:])[:, 0]  # noqa
    fa = f.reshape(P, R, n)[:, 0].abs()
    mix = lambda a, w: segment_sum(unit, w * a, U)  # noqa: E731
    av = mix(vals(avg), f.reshape(P, R, n)[:, 0][..., None])
    ea = mix(e_lin, fa[..., None])
    ext, sca, absb, back, scag = (av[..., k] for k in range(5))
    e_ext, e_sca, e_abs, e_back, e_scag = (ea[..., k] for k in range(5))

    def div(va, ea_, vb, eb):
        q = va / vb
        return (ea_ + q.abs() * eb) / vb.abs()

    zero = torch.zeros_like(ext)
    es = torch.stack([e_ext, e_sca, e_abs, e_back, div(sca, e_sca, ext, e_ext),
                      div(scag, e_scag, sca, e_sca), 4.0 * math.pi * div(ext, e_ext, back, e_back),
                      zero, zero, zero, zero], -1)
    cols = [es]
    if gr is not None:
        g_v = mix(vals(gr), f.reshape(P, R, n)[:, 0][..., None, None])
        g_e = mix(e_gr, fa[..., None, None])
        if normalized:
            a0, e0 = g_v[..., 0:1, 0:1], g_e[..., 0:1, 0:1]
            g_e = div(g_v, g_e, a0, e0)
        cols.append(g_e.reshape(g_e.shape[:2] + (-1,)))
    if ph is not None:
        p_v = mix(vals(ph), f.reshape(P, R, n)[:, 0][..., None, None])
        p_e = mix(e_ph, fa[..., None, None])
        if normalized:
            p_e = div(p_v, p_e, sca[..., None, None], e_sca[..., None, None])
        cols.append(p_e.reshape(p_e.shape[:2] + (-1,)))
    return torch.cat(cols, -1)


def _all_orders(units, wl, rtol, max_panels, refine_backscatter, allow_unvalidated, keys=(),
                shared: bool = False):
    """The reference's every order (GreekOrder::All, engine.rs:2034-2037): twice the largest N
    at each (unit, wavelength)'s final nodes, from a pass 1 of its own, (U, n)."""
    import torch

    from . import _native, adaptive
    from .size import MONODISPERSE

    dev = torch.device("cuda", torch.cuda.current_device())
    n = wl.size
    with torch.no_grad():
        parts = _parts(units, _Leaves(units, n, 1, dev, False))
        tasks = []
        for (u, kind, p, m_re, m_im, fr, _, L, cj) in parts:
            p0 = p.cpu().numpy()
            m0 = (m_re + 1j * m_im).cpu().numpy()
            f0 = fr.cpu().numpy()
            theta, dirs = _refined_derivatives(keys if u == 0 else (), cj, L)
            for i in range(n):
                tasks.append(adaptive.Task(int(kind), p0[i].copy(), float(wl[i]), m0[i].copy(),
                                           f0[i].copy(), theta=theta, dirs=dirs))
    if shared and not keys:
        from .shared import pass1_groups

        pass1_groups(tasks, [list(range(i, len(tasks), n)) for i in range(n)], rtol,
                     max_panels, refine_backscatter, allow_unvalidated)
    else:
        adaptive.pass1(tasks, rtol, max_panels, refine_backscatter, allow_unvalidated)
    lib = _native.load()
    out = np.zeros((len(units), n), np.int64)
    for k, q in enumerate(tasks):
        u, i = parts[k // n][0], k % n
        if q.kind == MONODISPERSE:
            r = q.p[0]
        else:  # the last panel's outermost Kronrod node
            r = np.exp(0.5 * (q.a[-1] + q.b[-1]) + 0.5 * (q.b[-1] - q.a[-1]) * adaptive._X[-1])
        out[u, i] = max(out[u, i], 2 * int(lib.scatrs_default_n_max(2.0 * np.pi * r / q.wl)))
    return out


def _clip_orders(V, E, D, DE, orders, W, n_mu):
    """Zeros each (unit, wavelength)'s Greek coefficients above its own order (the reference
    pads every result to the longest with zeros, scatrs-py/src/ensemble.rs:396-403)."""
    l = np.arange(W)
    for u in range(V.shape[0]):
        for i in range(V.shape[1]):
            cut = l > orders[u, i]
            if not cut.any():
                continue
            for k in range(6):
                cols = _N + k * W + np.flatnonzero(cut)
                V[u, i, cols] = 0.0
                E[u, i, cols] = 0.0
                if D is not None and u == 0:
                    D[:, cols, i] = 0.0
                    DE[:, cols, i] = 0.0


def _mu(mu):
    mu = np.ascontiguousarray(np.asarray(mu, np.float64).reshape(-1))
    return mu if mu.size else None


def ensemble_compute(items, wavelengths, mixture: bool, greek, all_orders: bool, mu,
                     normalized: bool, derivatives, rtol: float, max_panels: int,
                     allow_unvalidated: bool, refine_backscatter: bool, *,
                     n_radius: int | None = None, n_angles: int | None = None,
                     adaptive: bool | None = None) -> dict:
    """The ensemble outputs of the components at each wavelength as the reference's
    ``_scatrs.ensemble_compute`` returns them (scatrs-py/src/ensemble.rs:290-355): by the
    adaptive engine unless ``n_radius`` is given (or ``adaptive`` is False), else on the
    fixed path (``n_radius`` radius nodes per part, 128 without it, the Greek coefficients
    projected on ``n_angles`` Gauss-Legendre cosines, 128 by default). The adaptive
    engine (scatrs_amd.adaptive: rtol, max_panels and refine_backscatter as the reference
    uses them; n_angles=None sizes each Greek projection's rule to its nodes' largest N;
    derivatives on the final panels with the range held fixed); ``derivatives`` are indices
    into ``ensemble_parameters(items, mixture)``."""
    comps = _components(items)
    wl = np.ascontiguousarray(np.asarray(wavelengths, np.float64).reshape(-1))
    mu = _mu(mu)
    if not comps:
        raise InvalidInputError("a mixture needs a component")
    names, keys = _parameters(comps, mixture)
    chosen = []
    for i in derivatives:
        if not (0 <= i < len(keys)):
            raise InvalidInputError(f"no parameter {i}")
        if keys[i] in chosen:
            raise InvalidInputError(f"parameter {names[i]} requested twice")
        chosen.append(keys[i])
    _validate(comps, wl.size)
    order = _order(greek, all_orders)
    n_mu = 0 if mu is None else mu.size
    _check_options(rtol, max_panels, wl.size, n_mu, order, len(chosen))
    if adaptive is None:
        adaptive = n_radius is None
    orders = None
    if order == "all":
        if not adaptive:
            raise InvalidInputError("every Greek order (all_orders) needs the adaptive engine; "
                                    "the fixed path takes greek=L")
        orders = _all_orders([comps], wl, rtol, max_panels, refine_backscatter,
                             allow_unvalidated, chosen)
        order = int(orders.max())
    W = 0 if order is None else order + 1
    if adaptive:
        V, E, D, DE, conv, panels = _run_adaptive(
            [comps], wl, order=order, mu=mu, normalized=normalized, keys=chosen, rtol=rtol,
            max_panels=max_panels, refine_backscatter=refine_backscatter, n_angles=n_angles,
            allow_unvalidated=allow_unvalidated)
        if orders is not None:
            _clip_orders(V, E, D, DE, orders, W, n_mu)
        out = _dict(V, E, D, DE, conv, np.zeros(1, np.int64), wl.size, W, n_mu)
        out["panels"] = panels.reshape(-1).astype(np.uint64)
        return out
    V, E, D, DE, conv, counts = _run(
        [comps], wl, order=order, mu=mu, normalized=normalized, keys=chosen, rtol=rtol,
        refine_backscatter=refine_backscatter, n_radius=128 if n_radius is None else n_radius,
        n_angles=128 if n_angles is None else n_angles, allow_unvalidated=allow_unvalidated)
    return _dict(V, E, D, DE, conv, counts, wl.size, W, n_mu)


def ensemble_compute_batch(specs, particle_spec, wavelengths, greek, all_orders: bool, mu,
                           normalized: bool, rtol: float, max_panels: int,
                           allow_unvalidated: bool, refine_backscatter: bool, *,
                           n_radius: int | None = None, n_angles: int | None = None,
                           adaptive: bool | None = None, shared_panels: bool = False) -> dict:
    """The outputs of each distribution (tagged tuples ``specs``) of one particle at each
    wavelength, without derivatives, one row per (distribution, wavelength),
    distribution-major, as the reference's ``_scatrs.ensemble_compute_batch``
    (scatrs-py/src/ensemble.rs:480-504); every distribution in the one library call.

    ``shared_panels`` (adaptive engine): at each wavelength the distributions share one panel
    set and each node is solved once for all of them, as the reference's batch does
    (engine.rs:257-296; scatrs_amd.shared). By default each distribution gets its own panels,
    with the results of its single call: on the GPU the node solves are cheap and the shared
    set's finer union panels cost more host work per distribution than they save."""
    m, fr = _particle(particle_spec)
    wl = np.ascontiguousarray(np.asarray(wavelengths, np.float64).reshape(-1))
    mu = _mu(mu)
    units = [[_Component(_distribution(s), m, fr, 1.0)] for s in specs]
    for u in units:
        _validate(u, wl.size)
    order = _order(greek, all_orders)
    n_mu = 0 if mu is None else mu.size
    _check_options(rtol, max_panels, len(units) * wl.size, n_mu, order, 0)
    if adaptive is None:
        adaptive = n_radius is None
    orders = None
    if order == "all":
        if not adaptive:
            raise InvalidInputError("every Greek order (all_orders) needs the adaptive engine; "
                                    "the fixed path takes greek=L")
        orders = _all_orders(units, wl, rtol, max_panels, refine_backscatter,
                             allow_unvalidated, shared=shared_panels) if units \
            else np.zeros((0, wl.size), np.int64)
        order = int(orders.max()) if orders.size else 0
    W = 0 if order is None else order + 1
    if not units:
        import torch

        z = torch.zeros((0, wl.size, _N + 6 * W + 6 * n_mu), dtype=torch.float64)
        return _dict(z, z, None, None, torch.zeros((0, wl.size), dtype=torch.bool),
                     np.zeros(0, np.int64), wl.size, W, n_mu)
    if adaptive:
        V, E, _, _, conv, panels = _run_adaptive(
            units, wl, order=order, mu=mu, normalized=normalized, keys=[], rtol=rtol,
            max_panels=max_panels, refine_backscatter=refine_backscatter, n_angles=n_angles,
            allow_unvalidated=allow_unvalidated, shared=shared_panels)
        if orders is not None:
            _clip_orders(V, E, None, None, orders, W, n_mu)
        out = _dict(V, E, None, None, conv, np.zeros(len(units), np.int64), wl.size, W, n_mu)
        out["panels"] = panels.reshape(-1).astype(np.uint64)
        return out
    V, E, _, _, conv, counts = _run(
        units, wl, order=order, mu=mu, normalized=normalized, keys=[], rtol=rtol,
        refine_backscatter=refine_backscatter, n_radius=128 if n_radius is None else n_radius,
        n_angles=128 if n_angles is None else n_angles, allow_unvalidated=allow_unvalidated)
    return _dict(V, E, None, None, conv, counts, wl.size, W, n_mu)
