"""Adaptive size integration: the reference's Gauss-Kronrod engine (scatrs/src/ensemble/
engine.rs:307-943,1423-1766, kronrod.rs:5-52, range.rs:47-93) on the GPU node evaluator
``scatrs_ensemble_nodes``.

Each task (one part of a distribution at one wavelength) is integrated over panels in t = ln r
with the 15-point Kronrod rule and its embedded 7-point Gauss rule. The host places the
panels; every evaluation is one batched library call over all open panels of all tasks.

* Range: the ln r interval outside which the area-weighted law r^2 n(r) holds at most rtol/10
  of its total, split evenly between the tails. Z and the moments <r^2>, <r^3> come from their
  own integration over the number tail below and the volume tail above (moment_integrals), as
  the reference's averages take them from its moment integration. The tails are found from the library's own
  log-density on a fine grid over three times the fixed path's range. A power law, a tabulated
  segment and a monodisperse law keep their supports.
* Pass 1: the refined quantities are the reference's (engine.rs:620-690): <C_ext>, <C_sca>,
  <C_sca g> and <C_abs>, plus <C_back> with ``refine_backscatter``. Each quantity's error is the
  sum over panels of |Kronrod - Gauss|. Its target is rtol times |C_ext| (C_ext and C_abs),
  |C_sca| (C_sca and C_sca g), or max(|C_back|, |C_sca|) (C_back). While a task misses a target:
  * the range grows by one panel of the outer panel's width on a side whose outer panel holds
    more than rtol/10 of a quantity (the reference's extension);
  * every panel whose error exceeds the target over the number of panels is bisected, and
    always the worst one.
  A task stops unconverged when the next bisection would pass ``max_panels``.
* Pass 2: every output (the scalars, the Greek coefficients, the phase matrix) is evaluated on
  the final panels' Kronrod nodes, and with the Gauss weights on the same nodes (the same
  call); the error estimate comes from the two rules' outputs. The panels go in rows of 16, grouped by the angle rule
  their Greek projection needs to be exact, N_max + (L+1)/2 cosines at the row's largest size
  (the reference's bands, engine.rs:945-1017), and the rows are summed by their weights.

The panels are this engine's own: the reference's exact panel decisions depend on its
evaluation order. Both engines meet rtol, so their results agree to about rtol, not bitwise.
"""

from __future__ import annotations

import ctypes as C
import re
from dataclasses import dataclass, field
from functools import lru_cache

import numpy as np

from . import _native
from .errors import InvalidInputError
from .size import MONODISPERSE, POWER_LAW, SEGMENT

# QUADPACK's qk15 (Piessens, de Doncker-Kapenga, Ueberhuber and Kahaner, 1983): the Kronrod
# abscissae in [0, 1] (descending), their weights, and the 7-point Gauss weights of the
# abscissae 1, 3, 5 and 7 of the list.
_XGK = np.array([0.991455371120812639206854697526329, 0.949107912342758524526189684047851,
                 0.864864423359769072789712788640926, 0.741531185599394439863864773280788,
                 0.586087235467691130294144845693013, 0.405845151377397166906606412076961,
                 0.207784955007898467600689403773245, 0.0])
_WGK = np.array([0.022935322010529224963732008058970, 0.063092092629978553290700663189204,
                 0.104790010322250183839876322541518, 0.140653259715525918745189590510238,
                 0.169004726639267902826583426598550, 0.190350578064785409913256402421014,
                 0.204432940075298892414161999234649, 0.209482141084727828012999174891714])
_WG = np.array([0.129484966168869693270611432679082, 0.279705391489276667901467771423780,
                0.381830050505118944950369775488975, 0.417959183673469387755102040816327])

_INITIAL_PANELS = 4
_MAX_ROUNDS = 64
_GRID = 1025


def kronrod15():
    """(x (15,) ascending in [-1, 1], Kronrod weights (15,), Gauss weights (15,), 0 at the
    Kronrod-only nodes)."""
    x = np.concatenate([-_XGK[:7], [0.0], _XGK[6::-1]])
    wk = np.concatenate([_WGK[:7], [_WGK[7]], _WGK[6::-1]])
    wg = np.zeros(15)
    for j, k in enumerate((1, 3, 5)):
        wg[k] = wg[14 - k] = _WG[j]
    wg[7] = _WG[3]
    return x, wk, wg


_X, _WK, _WGF = kronrod15()
_GAUSS = np.flatnonzero(_WGF)


def _dbl(a):
    return np.ascontiguousarray(a, np.float64).ctypes.data_as(C.POINTER(C.c_double))


def check_parameters(kind: int, p) -> None:
    """InvalidInputError with the reference's message for invalid distribution parameters
    (scatrs_size_message)."""
    q = np.ascontiguousarray(p, np.float64)
    buf = C.create_string_buffer(512)
    if _native.load().scatrs_size_message(int(kind), _dbl(q), buf, 512) != 0:
        raise InvalidInputError(buf.value.decode())


def size_range(kind: int, p) -> tuple[float, float]:
    """The fixed path's ln r range of a distribution (scatrs_size_range)."""
    q = np.ascontiguousarray(p, np.float64)
    lo, hi = C.c_double(), C.c_double()
    if _native.load().scatrs_size_range(int(kind), _dbl(q), C.byref(lo), C.byref(hi)) != 0:
        check_parameters(kind, q)
        raise InvalidInputError(f"invalid parameters {tuple(q)} of size distribution kind {kind}")
    return lo.value, hi.value


def log_density(kind, p, r, derivatives: bool = False):
    """ln u(r), the number density per unit radius up to a constant (scatrs_size_log_density):
    kind a scalar or one per radius, p (3,) or (n, 3), r (n,); with ``derivatives`` also
    d ln u / d p (n, 3)."""
    r = np.ascontiguousarray(r, np.float64).reshape(-1)
    n = r.size
    k = np.ascontiguousarray(np.broadcast_to(np.asarray(kind, np.int32), (n,)))
    q = np.ascontiguousarray(np.broadcast_to(np.asarray(p, np.float64).reshape(-1, 3), (n, 3)))
    out = np.empty_like(r)
    d = np.empty((n, 3)) if derivatives else None
    if _native.load().scatrs_size_log_density(
            n, k.ctypes.data_as(C.POINTER(C.c_int32)), _dbl(q), _dbl(r), _dbl(out),
            None if d is None else C.c_void_p(d.ctypes.data)) != 0:
        raise InvalidInputError("invalid parameters of a size distribution")
    return (out, d) if derivatives else out


def integration_range(kind: int, p, rtol: float) -> tuple[float, float, float]:
    """(t_lo, t_hi, c): the ln r range whose area-weighted tails hold at most rtol/10 of the
    area, and c, the largest ln of the density per ln r there (the weights' common scale)."""
    lo, hi = size_range(kind, p)
    if kind in (POWER_LAW, SEGMENT):
        t = np.linspace(lo, hi, 257)
        return lo, hi, float(np.max(log_density(kind, p, np.exp(t)) + t))
    mid, half = 0.5 * (lo + hi), 0.5 * (hi - lo)
    t = np.linspace(mid - 3.0 * half, mid + 3.0 * half, _GRID)
    a = log_density(kind, p, np.exp(t)) + 3.0 * t  # ln of r^2 n(r) per ln r
    f = np.exp(a - np.max(a))
    cum = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(t))])
    eps = 0.05 * rtol * cum[-1]  # rtol/10 in all, half per tail
    i_lo = max(int(np.searchsorted(cum, eps, side="right")) - 1, 0)
    i_hi = min(int(np.searchsorted(cum, cum[-1] - eps, side="left")), t.size - 1)
    t_lo, t_hi = t[i_lo], t[i_hi]
    if not t_hi > t_lo:
        t_lo, t_hi = lo, hi
    tt = np.linspace(t_lo, t_hi, 257)
    return t_lo, t_hi, float(np.max(log_density(kind, p, np.exp(tt)) + tt))


def _tail_cuts(kind, p, eps, k_lo: int, k_hi: int):
    """ln r cuts below which the law weighted by r^k_lo, and above which the law weighted by
    r^k_hi, hold at most eps of their totals (on a fine grid over three times the fixed
    path's range)."""
    lo, hi = size_range(kind, p)
    mid, half = 0.5 * (lo + hi), 0.5 * (hi - lo)
    t = np.linspace(mid - 3.0 * half, mid + 3.0 * half, _GRID)
    lu = log_density(kind, p, np.exp(t))
    out = []
    for k, upper in ((k_lo, False), (k_hi, True)):
        a = lu + (k + 1.0) * t
        f = np.exp(a - np.max(a))
        cum = np.concatenate([[0.0], np.cumsum(0.5 * (f[1:] + f[:-1]) * np.diff(t))])
        if upper:
            out.append(t[min(int(np.searchsorted(cum, cum[-1] * (1 - eps), side="left")),
                             t.size - 1)])
        else:
            out.append(t[max(int(np.searchsorted(cum, eps * cum[-1], side="right")) - 1, 0)])
    return out[0], out[1]


_MOMENT_NODES = 1024


@lru_cache(maxsize=4)
def _legendre(n: int):
    """numpy's n-point Gauss-Legendre rule (an eigenvalue problem: computed once)."""
    return np.polynomial.legendre.leggauss(n)


def moment_integrals(kind: int, p, c: float, rtol: float):
    """Z = int u dr and M_k = int u r^k dr (k = 2, 3) over the moment range, and their
    derivatives with respect to p (3, 3) (rows Z, M2, M3), the density scaled by exp(-c) as
    the nodes' weights are. The range runs from where the number tail below holds rtol/20 of
    the number to where the volume tail above holds rtol/20 of the volume, so that Z, <r^2>
    and <r^3> miss no more than the reference's moment integration (moments.rs:62-127); a
    Gauss-Legendre rule of 1024 nodes in ln r. A power law or a tabulated segment: its
    support."""
    if kind in (POWER_LAW, SEGMENT):
        t_lo, t_hi = size_range(kind, p)
    else:
        t_lo, t_hi = _tail_cuts(kind, p, 0.05 * rtol, 0, 3)
    x, w = _legendre(_MOMENT_NODES)
    half, mid = 0.5 * (t_hi - t_lo), 0.5 * (t_hi + t_lo)
    t = mid + half * x
    lu, d = log_density(kind, p, np.exp(t), derivatives=True)
    f = w * half * np.exp(lu + t - c)  # u r dt, scaled
    r = np.exp(t)
    vals = np.array([f.sum(), (f * r * r).sum(), (f * r ** 3).sum()])
    der = np.stack([(f[:, None] * d).sum(0), ((f * r * r)[:, None] * d).sum(0),
                    ((f * r ** 3)[:, None] * d).sum(0)])
    return vals, der


@dataclass
class Task:
    """One part at one wavelength: its distribution (kind, p), index and fractions, range and
    panels (edges a, b and the refined sums of each rule, NaN until evaluated), and its pass-1
    state."""

    kind: int
    p: np.ndarray
    wl: float
    m: np.ndarray  # (L,) complex
    fr: np.ndarray  # (L-1,)
    t_lo: float = 0.0
    t_hi: float = 0.0
    c: float = 0.0
    a: np.ndarray = field(default_factory=lambda: np.zeros(0))
    b: np.ndarray = field(default_factory=lambda: np.zeros(0))
    K: np.ndarray = field(default_factory=lambda: np.zeros((0, 5)))
    G: np.ndarray = field(default_factory=lambda: np.zeros((0, 5)))
    moments: np.ndarray = field(default_factory=lambda: np.zeros(3))  # Z, M2, M3
    d_moments: np.ndarray = field(default_factory=lambda: np.zeros((3, 3)))
    done: bool = False
    converged: bool = False
    # The derivatives pass 1 refines: the distribution parameters (indices into p) and the
    # particle's directions (2 l, 2 l + 1: Re m, Im m of layer l; 2 L + j: radius fraction j).
    theta: tuple = ()
    dirs: tuple = ()

    @property
    def layers(self) -> int:
        return self.m.size

    @property
    def n_panels(self) -> int:
        return max(self.a.size, 1)


class _TaskTable:
    """The per-task attributes the rounds of pass 1 read, as arrays indexed by task (built once
    per pass, after the ranges and scales are set): kind, p (T, 3), c, wavelength, layers,
    the number of refined directions."""

    def __init__(self, tasks):
        self.kind = np.array([q.kind for q in tasks], np.int32)
        self.p = np.array([q.p for q in tasks], np.float64).reshape(-1, 3)
        self.c = np.array([q.c for q in tasks], np.float64)
        self.wl = np.array([q.wl for q in tasks], np.float64)
        self.layers = np.array([q.m.size for q in tasks], np.int64)
        self.n_dirs = np.array([len(q.dirs) for q in tasks], np.int64)
        self.m0 = np.array([q.m[0] for q in tasks], np.complex128)


def _nodes(tasks, ti, a, b, idx, derivatives: bool = False, table: _TaskTable | None = None):
    """(r, W) (len(a), len(idx)) of the panels [a, b] (each of task ti) at the nodes idx of the
    rule, with the Kronrod weights at all 15 nodes or the Gauss weights at the 7; a panel of
    width 0 (padding) has weights 0. One density call for all panels. With ``derivatives``
    also d ln W / d p (len(a), len(idx), 3), the range held fixed."""
    x = _X[idx]
    w = _WK[idx] if idx.size == 15 else _WGF[idx]
    half, mid = 0.5 * (b - a), 0.5 * (a + b)
    t = mid[:, None] + half[:, None] * x[None, :]
    r = np.exp(t)
    ti = np.asarray(ti, np.int64)
    tab = table if table is not None else _TaskTable(tasks)
    kind = tab.kind[ti]
    p = tab.p[ti]
    c = tab.c[ti]
    k = idx.size
    out = log_density(np.repeat(kind, k), np.repeat(p, k, axis=0), r.reshape(-1), derivatives)
    lu = out[0] if derivatives else out
    u = np.exp(lu.reshape(r.shape) + t - c[:, None])
    if derivatives:
        return r, w[None, :] * half[:, None] * u, out[1].reshape(r.shape + (3,))
    return r, w[None, :] * half[:, None] * u


def _call(tasks, ti, nodes, *, order=None, n_angles=2, mu=None, want_phase=False,
          normalized=False, allow_unvalidated=False):
    """The rows (task index ti (R,), nodes (R, k, 2) of (r, W)) by one library call per layer
    count: (scalars (R, 11), greek, phase) in row order, the rows' sums of W, and the terms."""
    import torch

    from .ensemble import ensemble_nodes_device

    dev = torch.device("cuda", torch.cuda.current_device())
    R = ti.size
    layers_t = np.array([q.layers for q in tasks])
    layers = layers_t[ti]
    sc = np.empty((R, 11))
    gr = None if order is None else np.empty((R, 6, order + 1))
    n_phase = n_angles if mu is None else int(np.size(mu))
    ph = np.empty((R, 6, n_phase)) if want_phase else None
    z = nodes[..., 1].sum(1)
    terms = 0
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    wl_all = np.array([q.wl for q in tasks])
    for L in np.unique(layers):
        sel = np.flatnonzero(layers == L)
        tk = ti[sel]
        own = np.flatnonzero(layers_t == L)  # the tasks of L layers, then their rows
        pos = np.full(len(tasks), -1)
        pos[own] = np.arange(own.size)
        m = np.stack([tasks[i].m for i in own])[pos[tk]]  # (g, L) complex
        fr = np.stack([tasks[i].fr for i in own])[pos[tk]] if L > 1 else None
        s, g, p, nt = ensemble_nodes_device(
            t(nodes[sel]), t(wl_all[tk]), t(m[:, 0] if L == 1 else m),
            None if fr is None else t(fr), n_angles=n_angles, order=order,
            normalized=normalized, allow_unvalidated=allow_unvalidated, want_phase=want_phase,
            mu=mu)
        sc[sel] = s.cpu().numpy()
        if gr is not None:
            gr[sel] = g.cpu().numpy()
        if ph is not None:
            ph[sel] = p.cpu().numpy()
        terms += nt
    return sc, gr, ph, z, terms


def _refined(sc, z):
    """The refined sums [C_ext, C_sca, C_sca g, C_abs, C_back] of rows (normalized=False
    scalars, which are averages: times the row's sum of weights); 0 for an empty row."""
    s = np.stack([sc[:, 0], sc[:, 1], sc[:, 5] * sc[:, 1], sc[:, 2], sc[:, 3]], 1) * z[:, None]
    s[z == 0.0] = 0.0
    return s


def _prepare(tasks, rtol: float) -> None:
    """Every task's range, weight scale c and moment integrals, as integration_range and
    moment_integrals give them one by one, with one density call per step for all tasks."""
    live = [q for q in tasks if q.kind != MONODISPERSE]
    if not live:
        return
    # Ranges and moments depend on the law alone: once per distinct (kind, p), bit for bit
    # (a sweep over wavelengths shares them).
    key = np.concatenate([np.array([[q.kind] for q in live], np.float64),
                          np.stack([q.p for q in live])], 1)
    _, first, inv = np.unique(np.ascontiguousarray(key).view(np.dtype((np.void, 32))).ravel(),
                              return_index=True, return_inverse=True)
    T = first.size
    kind = key[first, 0].astype(np.int32)
    p = key[first, 1:]
    lo, hi = np.array([size_range(k, q) for k, q in zip(kind, p)]).T
    support = np.isin(kind, (POWER_LAW, SEGMENT))
    mid, half = 0.5 * (lo + hi), 0.5 * (hi - lo)
    span = np.where(support, 1.0, 3.0)
    g = np.linspace(-1.0, 1.0, _GRID)
    t = mid[:, None] + (span * half)[:, None] * g[None, :]
    lu = log_density(np.repeat(kind, _GRID), np.repeat(p, _GRID, axis=0),
                     np.exp(t).reshape(-1)).reshape(T, _GRID)
    dt = np.diff(t, axis=1)

    def cum(k):
        a = lu + (k + 1.0) * t
        f = np.exp(a - a.max(1, keepdims=True))
        c = np.concatenate([np.zeros((T, 1)), np.cumsum(0.5 * (f[:, 1:] + f[:, :-1]) * dt, 1)], 1)
        return c / c[:, -1:]

    eps = 0.05 * rtol
    rows = np.arange(T)

    def lower(c):
        return t[rows, np.maximum((c <= eps).sum(1) - 1, 0)]

    def upper(c):
        return t[rows, np.minimum(_GRID - (c >= 1.0 - eps).sum(1), _GRID - 1)]

    area = cum(2)
    t_lo = np.where(support, lo, lower(area))
    t_hi = np.where(support, hi, upper(area))
    bad = ~(t_hi > t_lo)
    t_lo, t_hi = np.where(bad, lo, t_lo), np.where(bad, hi, t_hi)
    m_lo = np.where(support, lo, lower(cum(0)))
    m_hi = np.where(support, hi, upper(cum(3)))
    inside = (t >= t_lo[:, None]) & (t <= t_hi[:, None])
    c = np.where(inside, lu + t, -np.inf).max(1)
    x, w = _legendre(_MOMENT_NODES)
    mh, mm = 0.5 * (m_hi - m_lo), 0.5 * (m_hi + m_lo)
    tm = mm[:, None] + mh[:, None] * x[None, :]
    r = np.exp(tm)
    lum, d = log_density(np.repeat(kind, _MOMENT_NODES), np.repeat(p, _MOMENT_NODES, axis=0),
                         r.reshape(-1), derivatives=True)
    f = w[None, :] * mh[:, None] * np.exp(lum.reshape(T, -1) + tm - c[:, None])
    d = d.reshape(T, _MOMENT_NODES, 3)
    fk = np.stack([f, f * r * r, f * r ** 3], 1)  # (T, 3, n)
    mom = fk.sum(2)
    dmom = np.einsum("tkn,tnp->tkp", fk, d)
    for q, i in zip(live, inv.reshape(-1)):
        q.t_lo, q.t_hi, q.c = float(t_lo[i]), float(t_hi[i]), float(c[i])
        q.moments, q.d_moments = mom[i].copy(), dmom[i].copy()


# The refined quantities of a block (engine.rs:27-51): C_ext, C_sca, C_sca g, C_abs, C_back, and
# the quantity whose scale each takes within its block (C_abs: C_ext's, the accuracy of
# C_ext - C_sca).
_Q = 5
_SCALE = np.array([0, 1, 2, 0, 4])
MIN_WIDTH = 1e-10  # panels narrower than this (in ln r) are not split (engine.rs:62)
_ELEMENT = re.compile(r"element (\d+): (.*)", re.S)


def _node_values(tasks, ti, r, allow_unvalidated: bool, table: _TaskTable | None = None):
    """The pass-1 quantities of single particles (engine.rs:379-465): at each node (radius r[i]
    of task ti[i]'s particle at its wavelength) the cross sections [C_ext, C_sca, C_sca g,
    C_abs, C_back] (n, 5), pi r^2 times the efficiencies, and their derivatives along each
    task's requested directions (n, D, 5) in the order of its ``dirs`` (D the most any of these
    tasks asks for, zeros past a task's own; None when none asks), and the terms. One library
    call per layer count: the sphere batch (scatrs_mie_efficiencies) for spheres,
    scatrs_layered_efficiencies for layered spheres. An invalid node raises the library's error
    with the reference's wavelength-and-radius context (engine.rs:124-131)."""
    import torch

    from . import layered, mie
    from .errors import OutsideValidatedDomainError

    dev = torch.device("cuda", torch.cuda.current_device())
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
    ti = np.asarray(ti, np.int64)
    tab = table if table is not None else _TaskTable(tasks)
    n = r.size
    D = int(tab.n_dirs[ti].max()) if ti.size else 0
    Q = np.empty((n, _Q))
    dQ = np.zeros((n, D, _Q)) if D else None
    wl = tab.wl[ti]
    x = 2.0 * np.pi * r / wl
    area = np.pi * r * r
    layers = tab.layers[ti]
    if not D and np.all(layers == 1):
        # Spheres without refined derivatives: the general loop below in one step.
        try:
            v, _, nt = mie.efficiencies_host(x, tab.m0[ti], allow_unvalidated=allow_unvalidated)
        except (InvalidInputError, OutsideValidatedDomainError) as e:
            hit = _ELEMENT.fullmatch(str(e))
            if hit is None:
                raise
            i = int(hit[1])
            raise type(e)(f"at wavelength {wl[i]:e}, radius {r[i]:e}: {hit[2]}") from None
        return (area * np.stack([v[0], v[1], v[1] * v[4], v[2], v[3]])).T, None, nt
    terms = 0
    for L in np.unique(layers):
        sel = np.flatnonzero(layers == L)
        tk = ti[sel]
        own = np.unique(tk)
        want = bool(tab.n_dirs[own].any())
        pos = np.full(len(tasks), -1)
        pos[own] = np.arange(own.size)
        m = np.stack([tasks[k].m for k in own])[pos[tk]]  # (g, L)
        try:
            if L == 1:
                v, jac, nt = mie.efficiencies_host(x[sel], m[:, 0], derivatives=want,
                                                   allow_unvalidated=allow_unvalidated)
            else:
                fr = np.stack([tasks[k].fr for k in own])[pos[tk]]  # (g, L-1)
                xl = np.concatenate([fr * x[sel, None], x[sel, None]], 1)
                v, jac, nt = layered.efficiencies_device(t(xl), t(m), derivatives=want,
                                                         allow_unvalidated=allow_unvalidated)
        except (InvalidInputError, OutsideValidatedDomainError) as e:
            hit = _ELEMENT.fullmatch(str(e))
            if hit is None:
                raise
            i = sel[int(hit[1])]
            raise type(e)(f"at wavelength {wl[i]:e}, radius {r[i]:e}: {hit[2]}") from None
        terms += nt
        if L != 1:
            v = v.cpu().numpy()  # q_ext, q_sca, q_abs, q_back, g
            jac = None if jac is None else jac.cpu().numpy()
        a = area[sel]
        Q[sel] = (a * np.stack([v[0], v[1], v[1] * v[4], v[2], v[3]])).T
        if not want:
            continue
        j = jac
        # every direction's Jacobian of (q_ext, q_sca, q_abs, q_back, g): (g, 2 L + L - 1, 5)
        if L == 1:
            dirs_all = np.stack([j[1], j[2]])
        else:
            ms = [j[l, 1 + h] for l in range(L) for h in (0, 1)]
            fs = [x[sel] * j[f, 0] for f in range(L - 1)]  # x_L d/dx_f at a fixed x_L
            dirs_all = np.stack(ms + fs)
        dirs_all = dirs_all.transpose(2, 0, 1)
        # the slots of each row's task
        idx = np.full((own.size, D), -1)
        for k, q in enumerate(own):
            idx[k, :len(tasks[q].dirs)] = tasks[q].dirs
        idx = idx[pos[tk]]  # (g, D)
        rows = np.arange(sel.size)[:, None]
        d = dirs_all[rows, np.maximum(idx, 0)]  # (g, D, 5)
        d = np.where(idx[..., None] >= 0, d, 0.0)
        sca, g = v[1][:, None], v[4][:, None]
        dQ[sel] = a[:, None, None] * np.stack(
            [d[..., 0], d[..., 1], d[..., 1] * g + sca * d[..., 4], d[..., 2], d[..., 3]], -1)
    return Q, dQ, terms


# Pass-1 rounds of at least this many sphere nodes, without refined derivatives, form their
# panels' sums on the device (_panel_sums_device).
_DEVICE_SUMS_MIN = 65536


def _panel_sums_device(ti, r, wk, fold, allow_unvalidated: bool, table: _TaskTable):
    """A pass-1 round's Kronrod sums, Gauss sums and Kronrod sums of |integrand| (n, 5) of the
    refined quantities of n panels of spheres (tasks ti, nodes r (n, 15), Kronrod weights wk
    (n, 15), the Gauss rule's fold), and the terms: the sphere batch solves the nodes on the
    device (scatrs_mie_efficiencies) and the cross sections and weighted sums are formed there,
    so only the sums come back. The same quantities as _node_values and the host sums, summed
    in the device's order. An invalid node raises with the wavelength-and-radius context."""
    import torch

    from . import mie
    from .errors import OutsideValidatedDomainError

    dev = torch.device("cuda", torch.cuda.current_device())
    n = ti.size
    rr = np.ascontiguousarray(r, np.float64).reshape(-1)
    # Only the radii, the weights and the panels' tasks go to the device; each node's x and m
    # come from the tasks' wavelengths and indices there (the host's operations, bit for bit).
    tn = torch.from_numpy(np.ascontiguousarray(ti, np.int64)).to(dev).repeat_interleave(15)
    rt = torch.from_numpy(rr).to(dev)
    x = 2.0 * np.pi * rt / torch.from_numpy(table.wl).to(dev)[tn]
    m = torch.from_numpy(table.m0).to(dev)[tn]
    try:
        v, _, nt = mie.efficiencies_device(x, m, allow_unvalidated=allow_unvalidated)
    except (InvalidInputError, OutsideValidatedDomainError) as e:
        hit = _ELEMENT.fullmatch(str(e))
        if hit is None:
            raise
        i = int(hit[1])
        raise type(e)(f"at wavelength {table.wl[ti[i // 15]]:e}, radius {rr[i]:e}: "
                      f"{hit[2]}") from None
    area = np.pi * rt * rt
    q = torch.stack([area * v[0], area * v[1], area * (v[1] * v[4]), area * v[2],
                     area * v[3]], 1).reshape(n, 15, _Q)
    w = torch.from_numpy(np.ascontiguousarray(wk)).to(dev)
    wg = w * torch.from_numpy(np.ascontiguousarray(fold)).to(dev)
    sums = torch.stack([(w[..., None] * q).sum(1), (wg[..., None] * q).sum(1),
                        (w[..., None] * q.abs()).sum(1)]).cpu().numpy()
    return sums[0], sums[1], sums[2], nt


def _targets(tot, asum, rtol: float, D: int, H: int, n_dir, n_theta, refine_backscatter: bool):
    """The targets (T, nq) and the end targets (T, nq) of every refined quantity from the
    tasks' Kronrod totals ``tot`` and their sums of |integrand| ``asum`` (engine.rs:620-716):
    the values to rtol of |C_ext| (C_ext, C_abs), |C_sca| (C_sca, C_sca g) or max(|C_back|,
    |C_sca|) (C_back); a derivative along the particle's directions, or a distribution
    parameter's derivative, to rtol of the integral of its absolute integrand (C_abs's: of
    C_ext's), inf for the blocks a task does not ask for; C_back in every block only with
    ``refine_backscatter``. An end is negligible below a tenth of its target, a distribution
    parameter's below a tenth of rtol times its own total, floored at 1e-3 of its absolute
    integral."""
    T = tot.shape[0]
    sca = np.abs(tot[:, 1])
    tgt = np.full(tot.shape, np.inf)
    tgt[:, :_Q] = rtol * np.stack([np.abs(tot[:, 0]), sca, sca, np.abs(tot[:, 0]),
                                   np.maximum(np.abs(tot[:, 4]), sca)], 1)
    end = tgt.copy()
    for blk in range(1, 1 + D + H):
        o = blk * _Q
        used = n_dir > blk - 1 if blk <= D else n_theta > blk - 1 - D
        scale = asum[:, o + _SCALE]
        t = np.where(used[:, None], rtol * scale, np.inf)
        tgt[:, o:o + _Q] = t
        if blk <= D:
            end[:, o:o + _Q] = t
        else:
            end[:, o:o + _Q] = np.where(used[:, None], rtol * np.maximum(
                np.abs(tot[:, o:o + _Q]), 1e-3 * scale), np.inf)
    if not refine_backscatter:
        tgt[:, _Q - 1::_Q] = np.inf
        end[:, _Q - 1::_Q] = np.inf
    assert tgt.shape == (T, _Q * (1 + D + H))
    return tgt, end


def pass1(tasks, rtol: float, max_panels: int, refine_backscatter: bool,
          allow_unvalidated: bool) -> int:
    """Places every task's panels (see the module's notes); returns the Mie terms evaluated.

    The panels of all tasks are one table sorted by (task, ln r); every round's decisions are
    segment reductions over it. Each panel's 15 nodes are solved once (the Gauss rule takes 7
    of them); its sums per refined quantity by both rules, and the Kronrod sums of the
    absolute integrands, are formed on the host (engine.rs:482-566)."""
    terms = 0
    _prepare(tasks, rtol)
    table = _TaskTable(tasks)
    live = np.array([q.kind != MONODISPERSE for q in tasks])
    for q in tasks:
        if q.kind == MONODISPERSE:
            q.done = q.converged = True
    ids = np.flatnonzero(live)
    T = len(tasks)
    if not ids.size:
        return terms
    kinds = np.array([q.kind for q in tasks])
    fixed_support = np.isin(kinds, (POWER_LAW, SEGMENT))
    n_dir = np.array([len(q.dirs) if lv else 0 for q, lv in zip(tasks, live)])
    n_theta = np.array([len(q.theta) if lv else 0 for q, lv in zip(tasks, live)])
    D, H = int(n_dir.max()), int(n_theta.max())
    theta = np.full((T, max(H, 1)), -1)
    for i in ids:
        theta[i, :n_theta[i]] = tasks[i].theta
    nq = _Q * (1 + D + H)
    t_lo = np.array([q.t_lo for q in tasks])
    t_hi = np.array([q.t_hi for q in tasks])
    g = np.linspace(0.0, 1.0, _INITIAL_PANELS + 1)
    edges = t_lo[ids, None] + (t_hi - t_lo)[ids, None] * g[None, :]
    task = np.repeat(ids, _INITIAL_PANELS)
    a = edges[:, :-1].reshape(-1)
    b = edges[:, 1:].reshape(-1)
    K, G, A = (np.full((a.size, nq), np.nan) for _ in range(3))
    done = ~live
    conv = ~live
    fold = np.where(_WGF > 0, _WGF / _WK, 0.0)[None, :]
    finished = []
    for rnd in range(_MAX_ROUNDS):
        pending = np.flatnonzero(np.isnan(K[:, 0]) & ~done[task])
        if pending.size:
            ti = task[pending]
            n = pending.size
            out = _nodes(tasks, ti, a[pending], b[pending], np.arange(15), derivatives=H > 0,
                         table=table)
            r, wk = out[0], out[1]
            if (D == 0 and H == 0 and n * 15 >= _DEVICE_SUMS_MIN
                    and np.all(table.layers[ti] == 1)):
                K[pending], G[pending], A[pending], nt = _panel_sums_device(
                    ti, r, wk, fold, allow_unvalidated, table)
                terms += nt
            else:
                q, dq, nt = _node_values(tasks, np.repeat(ti, 15), r.reshape(-1),
                                         allow_unvalidated, table=table)
                terms += nt
                q = q.reshape(n, 15, _Q)
                f = [q]
                if D:  # these panels' tasks may ask for fewer directions than the most asked
                    d = np.zeros((n * 15, D, _Q))
                    if dq is not None:
                        d[:, :dq.shape[1]] = dq
                    f.append(d.reshape(n, 15, D * _Q))
                for h in range(H):  # d u / d theta = u d ln u / d theta
                    k = theta[ti, h]
                    dl = out[2][np.arange(n)[:, None], np.arange(15)[None, :],
                                np.maximum(k, 0)[:, None]]
                    f.append(q * np.where(k >= 0, 1.0, 0.0)[:, None, None] * dl[..., None])
                f = np.concatenate(f, -1)  # (n, 15, nq): the integrands per unit weight
                K[pending] = np.einsum("nk,nkq->nq", wk, f)
                G[pending] = np.einsum("nk,nkq->nq", wk * fold, f)
                A[pending] = np.einsum("nk,nkq->nq", wk, np.abs(f))
        act = ~done[task]
        if not act.any():
            break
        # Per task: totals, error sums and targets (engine.rs:620-716).
        E = np.abs(K - G)
        ta = task[act]
        cols = (ta[:, None] * nq + np.arange(nq)[None, :]).reshape(-1)
        seg = lambda v: np.bincount(cols, v[act].reshape(-1),  # noqa: E731
                                    minlength=T * nq).reshape(T, nq)
        tot, esum, asum = seg(K), seg(E), seg(A)
        tgt, end_tgt = _targets(tot, asum, rtol, D, H, n_dir, n_theta, refine_backscatter)
        fail = esum > tgt
        n_pan = np.bincount(task, minlength=T)
        first = np.ones(task.size, bool)
        first[1:] = task[1:] != task[:-1]
        last = np.ones(task.size, bool)
        last[:-1] = task[1:] != task[:-1]
        edge_big = (np.abs(K) > 0.1 * end_tgt[task]).any(1) & ~fixed_support[task]
        grow_lo = np.zeros(T, bool)
        grow_hi = np.zeros(T, bool)
        grow_lo[task[first & edge_big]] = True
        grow_hi[task[last & edge_big]] = True
        has_fail = fail.any(1) & ~done
        wants = (has_fail | grow_lo | grow_hi) & ~done
        newly = ~wants & ~done
        conv |= newly
        done |= newly
        grow_lo &= ~done
        grow_hi &= ~done
        # Within max_panels: the range extensions first, then the panels whose error exceeds
        # their share of a missed target, worst first (engine.rs:881-909); none narrower than
        # MIN_WIDTH.
        room = max_panels - n_pan
        grow_lo &= room >= 1
        room = room - grow_lo
        grow_hi &= room >= 1
        room = room - grow_hi
        tg = np.where(tgt > 0, tgt, np.finfo(float).tiny)
        ratio = np.where(fail[task], E / tg[task], 0.0).max(1)
        cand = np.flatnonzero((ratio * n_pan[task] > 1.0) & ~done[task] & (b - a > MIN_WIDTH))
        order = cand[np.lexsort((-ratio[cand], task[cand]))]  # by task, worst first
        ot = task[order]
        rank = np.arange(order.size) - np.searchsorted(ot, ot)
        cut = np.zeros(task.size, bool)
        cut[order[rank < room[ot]]] = True
        n_cut = np.bincount(task[cut], minlength=T)
        stuck = wants & (n_cut == 0) & ~grow_lo & ~grow_hi  # out of panels: unconverged
        done |= stuck
        if not (cut.any() or grow_lo.any() or grow_hi.any()) or rnd == _MAX_ROUNDS - 1:
            break  # after the last round the open tasks stay unconverged, every panel evaluated
        fin = done[task]  # finished tasks' panels leave the table: later rounds skip them
        if fin.any():
            finished.append((a[fin], b[fin], task[fin], K[fin], G[fin], A[fin]))
            a, b, task, K, G, A, cut, first, last = (
                x[~fin] for x in (a, b, task, K, G, A, cut, first, last))
        mid = 0.5 * (a + b)
        keep = ~cut
        nc = int(cut.sum())
        na = [a[keep], a[cut], mid[cut]]
        nb = [b[keep], mid[cut], b[cut]]
        nt_ = [task[keep], task[cut], task[cut]]
        lo_t = np.flatnonzero(grow_lo)
        hi_t = np.flatnonzero(grow_hi)
        if lo_t.size:
            w = (b - a)[first][np.searchsorted(task[first], lo_t)]
            na.append(t_lo[lo_t] - w)
            nb.append(t_lo[lo_t].copy())
            nt_.append(lo_t)
            t_lo[lo_t] -= w
        if hi_t.size:
            w = (b - a)[last][np.searchsorted(task[last], hi_t)]
            na.append(t_hi[hi_t].copy())
            nb.append(t_hi[hi_t] + w)
            nt_.append(hi_t)
            t_hi[hi_t] += w
        fresh = np.full((2 * nc + lo_t.size + hi_t.size, nq), np.nan)
        K, G, A = (np.concatenate([v[keep], fresh]) for v in (K, G, A))
        a, b, task = np.concatenate(na), np.concatenate(nb), np.concatenate(nt_)
        order = np.lexsort((a, task))
        a, b, task, K, G, A = (v[order] for v in (a, b, task, K, G, A))
    if finished:
        a, b, task, K, G, A = (np.concatenate([f[j] for f in finished] + [x])
                               for j, x in enumerate((a, b, task, K, G, A)))
        order = np.lexsort((a, task))
        a, b, task, K, G = a[order], b[order], task[order], K[order], G[order]
    bounds = np.searchsorted(task, np.arange(T + 1))
    for i in ids:
        q = tasks[i]
        lo, hi = bounds[i], bounds[i + 1]
        q.a, q.b, q.K, q.G = a[lo:hi].copy(), b[lo:hi].copy(), K[lo:hi].copy(), G[lo:hi].copy()
        q.t_lo, q.t_hi = float(t_lo[i]), float(t_hi[i])
        q.converged = bool(conv[i])
        q.done = True
    return terms


_ROW_PANELS = 16  # panels per pass-2 row


def _ladder_sizes():
    out = [16]
    while out[-1] < 4096:
        out.append(int(np.ceil(out[-1] * 1.25)))
    return np.minimum(np.array(out), 4096)


_LADDER = _ladder_sizes()


def _rule_ladder(n):
    """An angle rule's size: the next of 16, 20, 25, ... (steps of 1.25x), at most 4096 (n a
    number or an array)."""
    i = np.minimum(np.searchsorted(_LADDER, n, side="left"), _LADDER.size - 1)
    return _LADDER[i]


def _rules_for(x_max, order, n_angles, single: bool):
    """The Greek projection's rule of pass-2 rows whose largest size parameter is x_max (an
    array): the caller's n_angles, or the smallest exact for the row's nodes, N_max + (order +
    1)/2 cosines, N_max = ceil(x + 6 x^(1/3) + 3) as the library's scatrs_default_n_max (on a
    ladder, so that rows share calls; one rule for all rows when ``single``, for the phase at the
    rule's cosines)."""
    x_max = np.asarray(x_max, np.float64)
    if order is None:
        return np.full(x_max.size, 2 if n_angles is None else int(n_angles))
    if n_angles is not None:
        return np.full(x_max.size, int(n_angles))
    n = np.ceil(x_max + 6.0 * np.cbrt(x_max) + 3.0).astype(np.int64)
    need = _rule_ladder(n + (order + 2) // 2).astype(np.int64)
    if single and need.size:
        need[:] = need.max()
    return need


def _row_rules(tasks, rows, order, n_angles, single: bool):
    """_rules_for the rows (task, first panel, end panel; first < 0: a monodisperse task's one
    node): each row's largest radius, its last panel's end (or the monodisperse radius)."""
    x = np.array([2.0 * np.pi * (tasks[ti].p[0] if lo < 0 else np.exp(tasks[ti].b[hi - 1])) /
                  tasks[ti].wl for ti, lo, hi in rows], np.float64)
    return _rules_for(x, order, n_angles, single)


def _panel_rows(tasks):
    """The pass-2 rows of every task with panels (not monodisperse), in task order, _ROW_PANELS
    panels each (the last ones shorter): each row's task and the ln r of its last panel's end."""
    sizes = np.array([q.a.size if q.kind != MONODISPERSE else 0 for q in tasks], np.int64)
    own = np.flatnonzero(sizes > 0)
    cnt = sizes[own]
    n_rows = -(-cnt // _ROW_PANELS)
    ti = np.repeat(own, n_rows)
    j = np.arange(int(n_rows.sum())) - np.repeat(np.cumsum(n_rows) - n_rows, n_rows)
    hi = np.minimum((j + 1) * _ROW_PANELS, sizes[ti])
    b_flat = np.concatenate([tasks[t].b for t in own]) if own.size else np.zeros(0)
    b_last = b_flat[np.repeat(np.cumsum(cnt) - cnt, n_rows) + hi - 1]
    return ti, b_last


def _averages(sc):
    """The linear averages [C_ext, C_sca, C_abs, C_back, C_sca g, r^2, r^3] of normalized=False
    scalars (numpy)."""
    return np.stack([sc[:, 0], sc[:, 1], sc[:, 2], sc[:, 3], sc[:, 5] * sc[:, 1],
                     sc[:, 9] / np.pi, sc[:, 10] * (3.0 / (4.0 * np.pi))], 1)


def _scalars(avg):
    """normalized=False scalars from the linear averages (the inverse of _averages)."""
    ext, sca, absb, back, scag, r2, r3 = avg.T
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.stack([ext, sca, absb, back, sca / ext, scag / sca, 4.0 * np.pi * ext / back,
                         np.zeros_like(ext), r3 / r2, np.pi * r2, 4.0 / 3.0 * np.pi * r3], 1)


def pass2(tasks, *, order, n_angles, mu, want_phase, allow_unvalidated):
    """Every output of every task on its final panels (normalized=False), by the Kronrod and
    by the Gauss rule: ((sc, gr, ph) Kronrod, (sc, gr, ph) Gauss, terms, the phase rule's size).

    A task's panels go in rows of 16 (the last padded with nodes of weight 0), so that every
    call has many rows of one size; rows are grouped by their Greek rule (the smallest exact
    one for their nodes unless the caller fixes n_angles), one call per rule. The rows'
    unnormalized outputs are linear in their sums, so a task's are the rows' weighted by their
    sums of weights."""
    rows = []
    for ti, q in enumerate(tasks):
        if q.kind == MONODISPERSE:
            rows.append((ti, -1, -1))
            continue
        for lo in range(0, q.a.size, _ROW_PANELS):
            rows.append((ti, lo, min(lo + _ROW_PANELS, q.a.size)))
    single = want_phase and mu is None
    rules = _row_rules(tasks, rows, order, n_angles, single)
    ti_rows = np.array([r[0] for r in rows])
    T = len(tasks)
    out = []
    terms = 0
    # Each row's panels, padded with panels of width 0 at its last edge (weights 0).
    R = len(rows)
    A = np.zeros((R, _ROW_PANELS))
    Bd = np.zeros((R, _ROW_PANELS))
    mono = np.zeros(R, bool)
    for j, (ti, lo, hi) in enumerate(rows):
        q = tasks[ti]
        if lo < 0:
            mono[j] = True
            A[j] = Bd[j] = np.log(q.p[0])
            continue
        n = hi - lo
        A[j, :n], Bd[j, :n] = q.a[lo:hi], q.b[lo:hi]
        A[j, n:] = Bd[j, n:] = q.b[hi - 1]
    # Both rules in one call per angle rule: the Gauss rows are the same 15 nodes per panel with
    # the Gauss weights (0 at the Kronrod-only nodes).
    r, wk = _nodes(tasks, np.repeat(ti_rows, _ROW_PANELS), A.reshape(-1), Bd.reshape(-1),
                   np.arange(15))
    fold = np.where(_WGF > 0, _WGF / _WK, 0.0)[None, :]
    both = []
    for w in (wk, wk * fold):
        nodes = np.stack([r.reshape(R, -1), w.reshape(R, -1)], -1)
        nodes[mono, :, 1] = 0.0
        nodes[mono, 0, 1] = 1.0  # a monodisperse law: its radius, weight 1
        both.append(nodes)
    n_phase = (int(rules[0]) if mu is None else int(np.size(mu))) if want_phase else 0
    acc = np.zeros((2, T, 7))
    zt = np.zeros((2, T))
    gr = None if order is None else np.zeros((2, T, 6, order + 1))
    ph = np.zeros((2, T, 6, n_phase)) if want_phase else None
    for na in np.unique(rules):
        sel = np.flatnonzero(rules == na)
        sc, g, p, _, nt = _call(tasks, np.concatenate([ti_rows[sel], ti_rows[sel]]),
                                np.concatenate([both[0][sel], both[1][sel]]), order=order,
                                n_angles=int(na), mu=mu, want_phase=want_phase,
                                normalized=False, allow_unvalidated=allow_unvalidated)
        terms += nt
        for k in range(2):
            rows_k = slice(k * sel.size, (k + 1) * sel.size)
            zs = both[k][sel][..., 1].sum(1)
            live = zs > 0.0
            t, zl = ti_rows[sel][live], zs[live]
            np.add.at(acc[k], t, zl[:, None] * _averages(sc[rows_k][live]))
            np.add.at(zt[k], t, zl)
            if gr is not None:
                np.add.at(gr[k], t, zl[:, None, None] * g[rows_k][live])
            if ph is not None:
                np.add.at(ph[k], t, zl[:, None, None] * p[rows_k][live])
    # The averages divide by the moment integration's Z and take its <r^2>, <r^3>
    # (engine.rs:1782-1836); a monodisperse law keeps its node's.
    mom = np.array([q.moments for q in tasks])
    single_node = np.array([q.kind == MONODISPERSE for q in tasks])
    out = []
    for k in range(2):
        zm = np.where(single_node, zt[k], mom[:, 0])
        a_k = acc[k]
        a_k[~single_node, 5] = mom[~single_node, 1]
        a_k[~single_node, 6] = mom[~single_node, 2]
        with np.errstate(divide="ignore", invalid="ignore"):
            zi = 1.0 / zm
        out.append((_scalars(a_k * zi[:, None]),
                    None if gr is None else gr[k] * zi[:, None, None],
                    None if ph is None else ph[k] * zi[:, None, None]))
    return out[0], out[1], terms, int(rules[0]) if len(rules) else 2


def _row_panels(tasks):
    """The pass-2 rows of every task with panels, in task order: the task of each row and the
    [a, b] of its up to _ROW_PANELS panels (nr, _ROW_PANELS), a short last row padded with
    panels of width 0 at the task's upper end."""
    own = np.array([t for t, q in enumerate(tasks) if q.kind != MONODISPERSE and q.a.size],
                   np.int64)
    if not own.size:
        return own, np.zeros((0, _ROW_PANELS)), np.zeros((0, _ROW_PANELS))
    cnt = np.array([tasks[t].a.size for t in own], np.int64)
    n_rows = -(-cnt // _ROW_PANELS)
    first_row = np.cumsum(n_rows) - n_rows
    j = np.arange(cnt.sum()) - np.repeat(np.cumsum(cnt) - cnt, cnt)  # panel within its task
    row = np.repeat(first_row, cnt) + j // _ROW_PANELS
    col = j % _ROW_PANELS
    pad = np.repeat(np.array([tasks[t].b[-1] for t in own]), n_rows)
    A = np.repeat(pad[:, None], _ROW_PANELS, 1)
    Bd = A.copy()
    A[row, col] = np.concatenate([tasks[t].a for t in own])
    Bd[row, col] = np.concatenate([tasks[t].b for t in own])
    return np.repeat(own, n_rows), A, Bd


def pass2_torch(tasks, params, *, R: int, idx, order, n_angles, mu, want_phase,
                allow_unvalidated, gauss_tasks=None, derivatives=None):
    """The Kronrod (idx all 15 nodes) or Gauss rule's pass 2, differentiable in every task's
    parameters (torch.autograd): params[t] = (p (R, 3), m_real (R, L), m_imag (R, L),
    fractions (R, L-1) or None), R replicas of task t's parameters. Returns the linear averages
    [C_ext, C_sca, C_abs, C_back, C_sca g, r^2, r^3] (T, R, 7), greek (T, R, 6, order+1) and
    phase (T, R, 6, n) (normalized=False; None when not asked for), the terms, the phase
    rule's size, and each row's sums (detached: key = task x R + replica, lin (.., 7), gr, ph,
    and z, every task-replica's Z), for the error bounds.

    The rows are pass2's. A row's sum of weights Z enters to first order in p, Z0 + (p - p0)
    . sum W d ln W / d p, which has the right value and derivative at p0; a monodisperse task is
    the fixed path's one node (its radius moves its node). With gauss_tasks (bool per task) and
    all 15 nodes, those tasks take the Gauss weights on the same nodes (0 at the Kronrod-only
    ones): a task list doubled this way gets both rules from one call per angle rule. The
    weights' derivatives d ln W / d p are formed when ``derivatives`` (default: when gradients
    are enabled; forward mode asks for them explicitly)."""
    import torch

    from .ensemble import (ensemble_fixed, ensemble_nodes, number_averages, segment_plan,
                           segment_sum)

    dev = params[0][0].device
    T = len(tasks)
    f64 = torch.float64
    Wd = 0 if order is None else order + 1
    single = want_phase and mu is None
    row_task, row_end = _panel_rows(tasks)
    mono = [ti for ti, q in enumerate(tasks) if q.kind == MONODISPERSE]
    wl_t = np.array([q.wl for q in tasks], np.float64)
    x_rows = np.concatenate([2.0 * np.pi * np.exp(row_end) / wl_t[row_task],
                             [2.0 * np.pi * tasks[t].p[0] / tasks[t].wl for t in mono]])
    rules = _rules_for(x_rows, order, n_angles, single)
    n_phase = (int(rules[0]) if mu is None else int(np.size(mu))) if want_phase else 0
    avg = torch.zeros((T * R, 7), dtype=f64, device=dev)
    zt = torch.zeros(T * R, dtype=f64, device=dev)
    gr = torch.zeros((T * R, 6, Wd), dtype=f64, device=dev) if order is not None else None
    ph = torch.zeros((T * R, 6, n_phase), dtype=f64, device=dev) if want_phase else None
    terms = 0
    rep = torch.arange(R, device=dev)

    def gather(tk, L):
        """The parameters of tasks tk (each replicated R times, task-major)."""
        pr = torch.cat([params[t][0] for t in tk])
        mr = torch.cat([params[t][1] for t in tk])
        mi = torch.cat([params[t][2] for t in tk])
        fr = torch.cat([params[t][3] for t in tk]) if L > 1 else None
        if L == 1:
            mr, mi = mr[:, 0], mi[:, 0]
        return pr, mr, mi, fr

    # Each row's sums (detached), at its place among all rows x replicas, for the error bounds.
    n_rows = len(row_task) + len(mono)
    row_lin = torch.zeros((n_rows * R, 7), dtype=f64, device=dev)
    row_gr = torch.zeros((n_rows * R, 6, Wd), dtype=f64, device=dev) if gr is not None else None
    row_ph = torch.zeros((n_rows * R, 6, n_phase), dtype=f64, device=dev) if ph is not None \
        else None
    row_key = torch.zeros(n_rows * R, dtype=torch.int64, device=dev)

    def add(key, z, sc, g, p_, place):
        """Adds rows with task-replica keys ``key`` (host array) at row places ``place``."""
        nonlocal avg, zt, gr, ph
        plan = segment_plan(key, dev)
        lin = z[:, None] * number_averages(sc)
        avg = avg + segment_sum(plan, lin, T * R)
        zt = zt + segment_sum(plan, z, T * R)
        row_lin[place] = lin.detach()
        row_key[place] = plan.index
        if gr is not None:
            gz = z[:, None, None] * g
            gr = gr + segment_sum(plan, gz, T * R)
            row_gr[place] = gz.detach()
        if ph is not None:
            pz = z[:, None, None] * p_
            ph = ph + segment_sum(plan, pz, T * R)
            row_ph[place] = pz.detach()

    wl_all = np.array([q.wl for q in tasks])
    layers = np.array([q.layers for q in tasks])
    nr = len(row_task)
    if nr:
        ti_rows, A, Bd = _row_panels(tasks)
        grad = torch.is_grad_enabled() if derivatives is None else bool(derivatives)
        out_n = _nodes(tasks, np.repeat(ti_rows, _ROW_PANELS), A.reshape(-1), Bd.reshape(-1),
                       idx, derivatives=grad)
        r, w = out_n[0], out_n[1]
        d = out_n[2] if grad else np.zeros(r.shape + (3,))
        if gauss_tasks is not None:
            fold = np.where(_WGF > 0, _WGF / _WK, 0.0)[None, :]
            g_rows = np.repeat(np.asarray(gauss_tasks)[ti_rows], _ROW_PANELS)
            w = np.where(g_rows[:, None], w * fold, w)
        nodes = np.stack([r.reshape(nr, -1), w.reshape(nr, -1)], -1)
        dl = d.reshape(nr, -1, 3)
        z0 = nodes[..., 1].sum(1)
        zd = (nodes[..., 1:2] * dl).sum(1)
        # The rows' arrays go to the device once; each group of rows (one angle rule and layer
        # count) takes its rows there, by one transfer of all the groups' row lists.
        groups = []
        for na in np.unique(rules[:nr]):
            for L in np.unique(layers[ti_rows]):
                sel = np.flatnonzero((rules[:nr] == na) & (layers[ti_rows] == L) & (z0 > 0.0))
                if sel.size:
                    groups.append((int(na), int(L), sel))
        if groups:
            to_dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
            nodes_d, z0_d = to_dev(nodes), to_dev(z0)
            dl_d, zd_d = (to_dev(dl), to_dev(zd)) if grad else (None, None)
            wl_d, ti_d = to_dev(wl_all), to_dev(ti_rows.astype(np.int64))
            sel_d = to_dev(np.concatenate([g_[2] for g_ in groups]).astype(np.int64))
        pos = 0
        for na, L, sel in groups:
            tk = ti_rows[sel]
            pr, mr, mi, fr = gather(tk, L)
            sd = sel_d[pos:pos + sel.size]
            pos += sel.size
            rep_rows = lambda a: a[sd].repeat_interleave(R, dim=0)  # noqa: E731
            tk_d = ti_d[sd]
            sc, g, p_ = ensemble_nodes(pr, mr, mi, fr, rep_rows(nodes_d),
                                       rep_rows(dl_d) if grad else None,
                                       wl_d[tk_d].repeat_interleave(R, dim=0), n_angles=na,
                                       order=order, normalized=False,
                                       allow_unvalidated=allow_unvalidated, mu=mu,
                                       phase=want_phase, replicas=R)
            z = rep_rows(z0_d)
            if grad:
                z = z + ((pr - pr.detach()) * rep_rows(zd_d)).sum(1)
            key = (tk[:, None] * R + np.arange(R)[None, :]).reshape(-1)
            place = (sd[:, None] * R + rep[None, :]).reshape(-1)
            add(key, z, sc, g, p_, place)
    for k, ti in enumerate(mono):
        na = int(rules[nr + k])
        L = tasks[ti].layers
        pr, mr, mi, fr = gather([ti], L)
        kind = torch.full((R,), MONODISPERSE, dtype=torch.int32, device=dev)
        sc, g, p_ = ensemble_fixed(kind, pr[:, 0], pr[:, 1], torch.full(
            (R,), tasks[ti].wl, dtype=f64, device=dev), mr, mi, fr, p2=pr[:, 2], n_radius=1,
            n_angles=na, order=order, normalized=False, fixed_range=True,
            allow_unvalidated=allow_unvalidated, mu=mu, phase=want_phase)
        add(ti * R + np.arange(R), torch.ones(R, dtype=f64, device=dev), sc, g, p_,
            (nr + k) * R + rep)
    # The moment integration's Z, M2, M3 (first order in p: the right value and derivative at
    # p0), in place of the nodes' own; a monodisperse law keeps its node's.
    pall = torch.cat([params[t][0] for t in range(T)])  # (T R, 3)
    dp = pall - pall.detach()
    mom = torch.from_numpy(np.repeat(np.array([q.moments for q in tasks]), R, axis=0)).to(dev)
    dmom = torch.from_numpy(np.repeat(np.array([q.d_moments for q in tasks]), R, axis=0)).to(dev)
    mom = mom + (dmom * dp[:, None, :]).sum(-1)
    single_node = torch.from_numpy(np.repeat(np.array([q.kind == MONODISPERSE for q in tasks]),
                                             R)).to(dev)
    zm = torch.where(single_node, zt, mom[:, 0])
    avg = torch.cat([avg[:, :5], torch.where(single_node[:, None], avg[:, 5:], mom[:, 1:])], 1)
    avg = avg / zm[:, None]
    gr = None if gr is None else gr / zm[:, None, None]
    ph = None if ph is None else ph / zm[:, None, None]
    shape = lambda a: None if a is None else a.reshape((T, R) + a.shape[1:])  # noqa: E731
    rows_out = dict(key=row_key, lin=row_lin, gr=row_gr, ph=row_ph, z=zm.detach(),
                    n_rows=len(row_task), n_mono=len(mono))
    return (shape(avg), shape(gr), shape(ph), terms, (int(rules[0]) if len(rules) else 2),
            rows_out)


def pass2_from_panels(tasks):
    """Pass 2 of a call that asks for neither Greek coefficients nor the phase matrix nor
    derivatives, from pass 1's own sums: pass 1 has already summed every final panel's refined
    quantities [C_ext, C_sca, C_sca g, C_abs, C_back] over the same nodes with the same
    weights, by the Kronrod (K) and the Gauss rule (G), which is all that pass 2 would compute
    again. Returns [(avg (T, 1, 7), None, None, 0, 2, rows)] for the Kronrod and the Gauss rule
    as pass2_torch returns one rule (R = 1): the linear averages over the moment integration's
    Z (r^2 and r^3 from it), and the rows' sums (rows of _ROW_PANELS panels, in task order) for
    the error bounds. Every task has panels (no monodisperse part)."""
    import torch

    dev = torch.device("cuda", torch.cuda.current_device())
    T = len(tasks)
    mom = np.array([q.moments for q in tasks], np.float64)  # (T, 3): Z, M2, M3
    row_task, _ = _panel_rows(tasks)
    out = []
    for sums in ("K", "G"):
        # [C_ext, C_sca, C_abs, C_back, C_sca g] from pass 1's [C_ext, C_sca, C_sca g, C_abs, C_back]
        lin = [getattr(q, sums)[:, [0, 1, 3, 4, 2]] for q in tasks]
        tot = np.array([x.sum(0) for x in lin])  # (T, 5)
        avg = np.concatenate([tot, mom[:, 1:]], 1) / mom[:, :1]
        rows = np.concatenate([np.add.reduceat(x, np.arange(0, x.shape[0], _ROW_PANELS), 0)
                               for x in lin], 0)
        rows = np.concatenate([rows, np.zeros((rows.shape[0], 2))], 1)
        to = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)  # noqa: E731
        d = {"key": to(row_task.astype(np.int64)), "lin": to(rows), "gr": None, "ph": None,
             "z": to(mom[:, 0])}
        out.append((to(avg).reshape(T, 1, 7), None, None, 0, 2, d))
    return out


def parts_adaptive(parts, wl: np.ndarray, *, order, n_angles, rtol: float, max_panels: int,
                   allow_unvalidated: bool, want_phase: bool, mu=None,
                   refine_backscatter: bool = False):
    """Each part (kind, p0, p1, p2, particle) at every wavelength by the adaptive engine with
    normalized=False, as _parts_device returns them (tensors (parts, wavelengths, ...)), with
    the Gauss rule's outputs, converged and panels (parts, wavelengths) and the ranges
    (parts, wavelengths, 2) in r: (K outputs, G outputs, terms, converged, panels, ranges)."""
    import torch

    if not (1e-13 <= rtol <= 0.1):
        raise InvalidInputError(f"rtol must be in [1e-13, 0.1], got {rtol}")
    if max_panels < 1:
        raise InvalidInputError("max_panels must be at least 1")
    nw = wl.size
    tasks = []
    for kind, p0, p1, p2, particle in parts:
        m = particle.indices(wl).reshape(particle.layers, -1)  # (L, nw)
        fr = np.asarray(particle.fractions(), np.float64)
        for i in range(nw):
            tasks.append(Task(int(kind), np.array([p0, p1, p2], np.float64),
                              float(wl.reshape(-1)[i]), m[:, i].copy(), fr))
    terms = pass1(tasks, rtol, max_panels, refine_backscatter, allow_unvalidated)
    K, G, t2, na = pass2(tasks, order=order, n_angles=n_angles, mu=mu, want_phase=want_phase,
                         allow_unvalidated=allow_unvalidated)
    dev = torch.device("cuda", torch.cuda.current_device())
    P = len(parts)

    def shape(out):
        sc, gr, ph = out
        f = lambda a: None if a is None else torch.from_numpy(  # noqa: E731
            a.reshape((P, nw) + a.shape[1:])).to(dev)
        return f(sc), f(gr), f(ph)

    conv = np.array([q.converged for q in tasks]).reshape(P, nw)
    panels = np.array([q.n_panels for q in tasks]).reshape(P, nw)
    ranges = np.array([(q.p[0], q.p[0]) if q.kind == MONODISPERSE
                       else (np.exp(q.t_lo), np.exp(q.t_hi)) for q in tasks]).reshape(P, nw, 2)
    return shape(K), shape(G), terms + t2, conv, panels, ranges, na
